"""tsdb database files and appends on the GPU: a saved set is byte for byte
the file the host writer makes from the same trajectories (the device's
summaries equal the host's); a loaded set, and a set grown by appends (with
and without its device arrays growing), give the same k-NN, bit for bit, as
a set created from the arrays, and the same values as the oracle; a file
whose stored MBRs disagree with its points is refused."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from helpers import planted, say, uniform_lengths, walks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    ts.init(0)
    L = uniform_lengths(61, 2500, 16, 300)
    L[::97] = 1
    db_xy, db_off = walks(61, L, 2e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 2500, 250), L[0:2500:250], 1e-3, 62)
    return db_xy, db_off, q_xy, q_off


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int32), b[1].view(np.int32))


PARAMS = [("dtw", {}), ("dtw", {"band": -10}), ("lcss", {"epsilon": 2e-3, "delta": 16}), ("frechet", {}),
          ("hausdorff", {})]


def test_save_equals_host_writer(tmp_path, data):
    db_xy, db_off = data[0], data[1]
    say("tsdb save")
    s = ts.TrajSet(db_xy, db_off)
    s.save(str(tmp_path / "dev.tsdb"))
    ts.db_write(str(tmp_path / "host.tsdb"), db_xy, db_off)
    assert open(tmp_path / "dev.tsdb", "rb").read() == open(tmp_path / "host.tsdb", "rb").read()


def test_loaded_set_gives_the_same_knn(tmp_path, data):
    db_xy, db_off, q_xy, q_off = data
    say("tsdb load")
    path = str(tmp_path / "db.tsdb")
    ts.db_write(path, db_xy, db_off)
    a, b = ts.TrajSet(db_xy, db_off), ts.TrajSet.load(path)
    assert np.array_equal(b.lengths, np.diff(db_off)) and b.info() == a.info()
    qs = ts.TrajSet(q_xy, q_off)
    eng = ts.Engine()
    for measure, kw in PARAMS:
        p = ts.params(measure, mode=ts.EXACT, k=10, **kw)
        got, want = eng.knn(b, qs, p), eng.knn(a, qs, p)
        assert same(got, want), (measure, kw)
    p = ts.params("dtw", mode=ts.EXACT, k=10)
    o = oracle.knn(oracle.params(oracle.DTW, oracle.EXACT, k=10), db_xy, db_off, q_xy, q_off)
    got = eng.knn(b, qs, p)
    assert np.array_equal(got[0], o[0]) and np.array_equal(got[1].view(np.int32), o[1].view(np.int32))


@pytest.mark.parametrize("grow", [True, False])
def test_appended_set_gives_the_same_knn(data, grow):
    db_xy, db_off, q_xy, q_off = data
    say("tsdb append", grow)
    cuts = [0, 700, 701, 1900, 2500]
    if grow:
        s = ts.TrajSet(db_xy[:db_off[cuts[1]]], db_off[:cuts[1] + 1])
    else:
        # room for everything: created full, reloaded with the first part
        s = ts.TrajSet(db_xy, db_off)
        s.reload(db_xy[:db_off[cuts[1]]], db_off[:cuts[1] + 1])
    for a, b in zip(cuts[1:-1], cuts[2:]):
        s.append(db_xy[db_off[a]:db_off[b]], db_off[a:b + 1] - db_off[a])
    assert np.array_equal(s.lengths, np.diff(db_off)) and s.count == 2500
    whole = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    eng = ts.Engine()
    for measure, kw in PARAMS:
        p = ts.params(measure, mode=ts.EXACT, k=10, **kw)
        assert same(eng.knn(s, qs, p), eng.knn(whole, qs, p)), (measure, kw)


def test_file_with_wrong_summaries_is_refused(tmp_path, data):
    db_xy, db_off = data[0], data[1]
    say("tsdb bad summaries")
    path = str(tmp_path / "db.tsdb")
    ts.db_write(path, db_xy, db_off)
    raw = bytearray(open(path, "rb").read())
    mbr_at = int(np.frombuffer(raw[48 + 3 * 8:48 + 4 * 8], "<u8")[0])  # section 3: MBRs
    x = np.frombuffer(raw[mbr_at + 16 * 5:mbr_at + 16 * 5 + 4], "<f4")[0]
    raw[mbr_at + 16 * 5:mbr_at + 16 * 5 + 4] = np.float32(x - 1.0).tobytes()
    open(path, "wb").write(bytes(raw))
    with pytest.raises(ts.TrajsimError, match="disagree with the points"):
        ts.TrajSet.load(path)
