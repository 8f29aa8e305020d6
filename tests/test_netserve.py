"""CPU checks of the socket front end (cudaruntime1_amd/netserve.py) with a
stand-in backend: frames round-trip, pipelined requests on several
connections come back matched by id, a refused request returns its status,
a malformed frame closes only its own connection, and unset parameters
travel as 'default'. The GPU test (test_gpu_netserve.py) serves a real
database."""
import socket
import threading

import numpy as np
import pytest

from cudaruntime1_amd import netserve as ns


class Busy(Exception):
    status = 6


class FakeBackend:
    """Result of a request: indices len(points) + 0..k-1, values the sum of
    its coordinates plus 0..k-1; refuses queries whose first x is negative."""

    def __init__(self):
        self.lock = threading.Lock()
        self.pending = {}
        self.seen = []
        self.next = 0

    def submit(self, points, prm):
        if points[0, 0] < 0:
            raise Busy("admission refused")
        with self.lock:
            self.next += 1
            self.pending[self.next] = (points.copy(), prm)
            self.seen.append(prm)
            return self.next

    def wait(self, ticket, k):
        with self.lock:
            pts, prm = self.pending.pop(ticket)
        base = float(pts.astype(np.float64).sum())
        return (np.arange(k, dtype=np.int32) + len(pts), (base + np.arange(k)).astype(np.float32), 0.25)


@pytest.fixture
def served():
    be = FakeBackend()
    srv = ns.NetServer(be).start()
    yield srv, be
    srv.stop()


def test_frames_round_trip():
    pts = np.array([[1.5, -2.0], [3.0, 4.25], [0.0, 1e-3]], np.float32)
    f = ns.encode_request(7, pts, **ns._fill({"measure": "lcss", "k": 3, "epsilon": 2e-3, "delta": 16}))
    rid, prm, got = ns.decode_request(f[:ns.REQ_HEAD.size], f[ns.REQ_HEAD.size:])
    assert rid == 7 and np.array_equal(got, pts)
    assert prm["measure"] == 1 and prm["k"] == 3 and prm["delta"] == 16 and prm["band"] == ns.DEFAULT
    assert prm["epsilon"] == np.float32(2e-3)
    r = ns.encode_response(9, 0, np.array([4, 5], np.int32), np.array([0.5, 1.5], np.float32), 2.5)
    rid, st, idx, val, lat, _ = ns.decode_response(r[:ns.RESP_HEAD.size], r[ns.RESP_HEAD.size:])
    assert (rid, st, lat) == (9, 0, 2.5) and idx.tolist() == [4, 5] and val.tolist() == [0.5, 1.5]
    with pytest.raises(ValueError):
        ns.decode_request(b"XXXX" + f[4:ns.REQ_HEAD.size], f[ns.REQ_HEAD.size:])


def test_pipelined_requests_on_several_connections(served):
    srv, be = served
    rng = np.random.default_rng(1)
    out = {}

    def run(c):
        cl = ns.Client(port=srv.port)
        reqs = [(rng.random((int(n), 2)).astype(np.float32), {"measure": "dtw", "k": 1 + int(n) % 7})
                for n in rng.integers(1, 40, 50)]
        out[c] = (reqs, cl.knn_many(reqs))
        cl.close()

    ts = [threading.Thread(target=run, args=(c,)) for c in range(4)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    for reqs, res in out.values():
        for (pts, kw), (idx, val, lat) in zip(reqs, res):
            k = kw["k"]
            assert idx.tolist() == list(range(len(pts), len(pts) + k))
            assert np.array_equal(val, (float(pts.astype(np.float64).sum()) + np.arange(k)).astype(np.float32))
            assert lat == 0.25


def test_refused_request_returns_its_status(served):
    srv, _ = served
    cl = ns.Client(port=srv.port)
    ok = np.ones((5, 2), np.float32)
    res = cl.knn_many([(ok, {}), (-ok, {}), (ok, {"k": 2})])
    assert res[0][0].tolist() == list(range(5, 15)) and res[2][0].tolist() == [5, 6]
    assert isinstance(res[1], ns.RequestError) and res[1].status == 6
    with pytest.raises(ns.RequestError):
        cl.knn(-ok)
    cl.close()


def test_malformed_frame_closes_only_its_connection(served):
    srv, _ = served
    bad = socket.create_connection(("127.0.0.1", srv.port), timeout=10)
    bad.sendall(b"JUNK" + bytes(60))
    assert bad.recv(16) == b""  # closed by the server
    bad.close()
    cl = ns.Client(port=srv.port)
    idx, _, _ = cl.knn(np.ones((3, 2), np.float32), k=4)
    assert idx.tolist() == [3, 4, 5, 6]
    cl.close()


def test_defaults_travel_as_default(served):
    srv, be = served
    cl = ns.Client(port=srv.port)
    cl.knn(np.ones((2, 2), np.float32), measure="euclidean")
    cl.knn(np.ones((2, 2), np.float32), measure="dtw", band=-10, epsilon=0.25)
    cl.close()
    a, b = be.seen[-2], be.seen[-1]
    assert a["measure"] == 4 and a["kc"] == ns.DEFAULT and a["band"] == ns.DEFAULT and np.isnan(a["epsilon"])
    assert b["band"] == -10 and b["epsilon"] == 0.25 and b["delta"] == ns.DEFAULT
