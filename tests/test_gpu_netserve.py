"""The socket front end over a real database: pipelined requests of every
measure and k in {1, 10, 100} give the engine's own top-k bit for bit, and the
command-line server (python -m cudaruntime1_amd.netserve) serves a tsdb file."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cudaruntime1_amd as ts
from cudaruntime1_amd import netserve as ns
from helpers import planted, say, uniform_lengths, walks

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def data():
    ts.init(0)
    L = uniform_lengths(71, 3000, 16, 200)
    db_xy, db_off = walks(71, L, 2e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 3000, 300), L[0:3000:300], 1e-3, 72)
    return db_xy, db_off, q_xy, q_off


REQS = [{"measure": "dtw", "k": 10}, {"measure": "dtw", "k": 1, "band": -10}, {"measure": "frechet", "k": 100},
        {"measure": "lcss", "k": 10, "epsilon": 2e-3, "delta": 16}, {"measure": "hausdorff", "k": 10}]


def want(eng, db, q, kw):
    p = ts.params(kw["measure"], mode=ts.EXACT, k=kw["k"], band=kw.get("band"), delta=kw.get("delta"),
                  epsilon=kw.get("epsilon"))
    qs = ts.TrajSet(q, np.array([0, len(q)], np.int64))
    idx, val, _ = eng.knn(db, qs, p)
    return idx[0], val[0]


def test_served_results_equal_engine(data):
    db_xy, db_off, q_xy, q_off = data
    say("netserve")
    db = ts.TrajSet(db_xy, db_off)
    srv = ns.NetServer(ns.EngineBackend(db, streams=2)).start()
    try:
        cl = ns.Client(port=srv.port)
        reqs = [(q_xy[q_off[i]:q_off[i + 1]], kw) for i in range(len(q_off) - 1) for kw in REQS]
        got = cl.knn_many(reqs)
        cl.close()
        eng = ts.Engine()
        for (q, kw), r in zip(reqs, got):
            idx, val = want(eng, db, q, kw)
            assert np.array_equal(r[0], idx) and np.array_equal(r[1].view(np.int32), val.view(np.int32)), kw
    finally:
        srv.stop()
        srv.backend.close()


def test_command_line_server_serves_a_database_file(tmp_path, data):
    db_xy, db_off, q_xy, q_off = data
    say("netserve cli")
    path = str(tmp_path / "db.tsdb")
    ts.db_write(path, db_xy, db_off)
    proc = subprocess.Popen([sys.executable, "-u", "-m", "cudaruntime1_amd.netserve", "--db", path, "--port", "0"],
                            cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    try:
        line = proc.stdout.readline()
        assert line.startswith("serving 3000 trajectories on 127.0.0.1:"), line
        port = int(line.rsplit(":", 1)[1])
        cl = ns.Client(port=port)
        q = q_xy[q_off[0]:q_off[1]]
        idx, val, _ = cl.knn(q, measure="dtw", k=10)
        cl.close()
        w = want(ts.Engine(), ts.TrajSet(db_xy, db_off), q, {"measure": "dtw", "k": 10})
        assert np.array_equal(idx, w[0]) and np.array_equal(val.view(np.int32), w[1].view(np.int32))
    finally:
        proc.terminate()
        proc.wait(timeout=30)
