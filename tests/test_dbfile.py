"""CPU checks of the tsdb database file (csrc/dbfile.hip): a file written from
host arrays holds, section by section, what numpy computes from the same
trajectories (the device layout, MBRs, endpoints, the length index), and the
loader rejects damaged files before it touches a GPU. GPU round trips (save,
load, append against the same k-NN) are in test_gpu_dbfile.py."""
import os

import numpy as np
import pytest

import cudaruntime1_amd as ts

HDR = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"), ("count", "<i4"),
                ("min_len", "<i4"), ("max_len", "<i4"), ("pad", "<i4"), ("points", "<i8"),
                ("dev_points", "<i8"), ("sec_off", "<u8", 7), ("sec_bytes", "<u8", 7), ("checksum", "<u8"),
                ("reserved", "u1", 88)])
LENGTHS = [1, 2, 7, 8, 9, 15, 16, 17, 64, 65, 300, 1, 33, 1024, 5]


def fnv1a(b):
    h = 0xcbf29ce484222325
    for x in b:
        h = ((h ^ x) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def make(tmp_path, lengths=LENGTHS, seed=3):
    lens = np.array(lengths, np.int32)
    xy, off = ts.gen_walks(seed, lens, 1e-2)
    path = str(tmp_path / "db.tsdb")
    ts.db_write(path, xy, off)
    return path, xy, off


def sections(path):
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:256], HDR)[0]
    sec = []
    for i, dt in enumerate(["<i4", "<f4", "<f4", "<f4", "<f4", "<i4", "<i8"]):
        a, n = int(h["sec_off"][i]), int(h["sec_bytes"][i])
        sec.append(np.frombuffer(raw[a:a + n], dt))
    return raw, h, sec


def test_header_and_sections_match_numpy(tmp_path):
    path, xy, off = make(tmp_path)
    raw, h, (lens, x, y, mbr, ends, order, cls) = sections(path)
    L = np.diff(off).astype(np.int32)
    assert h["magic"] == b"TSDBSOA1" and h["version"] == 1 and h["header_bytes"] == 256 and h["pad"] == 8
    assert h["count"] == len(L) and h["min_len"] == L.min() and h["max_len"] == L.max()
    assert h["points"] == off[-1]
    hb = bytearray(raw[:256])
    hb[8 + 8 + 16 + 16 + 112:8 + 8 + 16 + 16 + 112 + 8] = b"\0" * 8
    assert fnv1a(bytes(hb)) == int(h["checksum"])
    assert all(int(o) % 256 == 0 for o in h["sec_off"])
    assert np.array_equal(lens, L)
    # padded SoA layout: trajectory i at the sum of earlier lengths rounded up to 8
    pad = (L + 7) // 8 * 8
    doff = np.concatenate([[0], np.cumsum(pad)])
    assert h["dev_points"] == doff[-1] and len(x) == doff[-1]
    for i in range(len(L)):
        t = xy[off[i]:off[i + 1]]
        seg_x, seg_y = x[doff[i]:doff[i + 1]], y[doff[i]:doff[i + 1]]
        assert np.array_equal(seg_x[:L[i]], t[:, 0]) and np.array_equal(seg_y[:L[i]], t[:, 1])
        assert np.all(seg_x[L[i]:] == t[-1, 0]) and np.all(seg_y[L[i]:] == t[-1, 1])
        want_mbr = [t[:, 0].min(), t[:, 0].max(), t[:, 1].min(), t[:, 1].max()]
        assert np.array_equal(mbr[4 * i:4 * i + 4], np.array(want_mbr, np.float32))
        assert np.array_equal(ends[4 * i:4 * i + 4], np.array([t[0, 0], t[0, 1], t[-1, 0], t[-1, 1]], np.float32))
    # length index: ids by (length, id), class c = lengths in [2^c, 2^(c+1))
    assert np.array_equal(order, np.argsort(L, kind="stable"))
    assert len(cls) == 26 and cls[0] == 0 and cls[25] == len(L)
    for c in range(25):
        sel = L[order[cls[c]:cls[c + 1]]]
        assert np.all((sel >= 2 ** c) & (sel < 2 ** (c + 1)))


def test_writer_is_deterministic(tmp_path):
    p1, _, _ = make(tmp_path)
    a = open(p1, "rb").read()
    os.rename(p1, str(tmp_path / "a.tsdb"))
    p2, _, _ = make(tmp_path)
    assert open(p2, "rb").read() == a


def damage(path, at, value):
    raw = bytearray(open(path, "rb").read())
    raw[at:at + len(value)] = value
    open(path, "wb").write(bytes(raw))


@pytest.mark.parametrize("how,msg", [("magic", "not a tsdb file"), ("version", "unsupported version"),
                                     ("count", "checksum"),
                                     ("truncate", "out of place"), ("length", "lengths disagree"),
                                     ("order", "length index"), ("short", "too short")])
def test_loader_rejects_damaged_files(tmp_path, how, msg):
    path, _, _ = make(tmp_path)
    raw, h, _ = sections(path)
    if how == "magic":
        damage(path, 0, b"XSDB")
    elif how == "version":
        damage(path, 8, (2).to_bytes(4, "little"))
    elif how == "count":
        damage(path, 16, (int(h["count"]) - 1).to_bytes(4, "little"))
    elif how == "truncate":
        open(path, "wb").write(raw[:int(h["sec_off"][2]) + 16])
    elif how == "length":
        damage(path, int(h["sec_off"][0]) + 4 * 3, (9).to_bytes(4, "little"))
    elif how == "order":
        o = int(h["sec_off"][5])
        a, b = raw[o:o + 4], raw[o + 4:o + 8]
        damage(path, o, b + a)
    else:
        open(path, "wb").write(raw[:100])
    with pytest.raises(ts.TrajsimError, match=msg):
        ts.TrajSet.load(path)


def test_writer_rejects_bad_lengths(tmp_path):
    xy = np.zeros((4, 2), np.float32)
    with pytest.raises(ts.TrajsimError):
        ts.db_write(str(tmp_path / "bad.tsdb"), xy, np.array([0, 2, 2, 4], np.int64))
