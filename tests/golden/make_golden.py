"""Writes the golden fixtures in this directory from the reference's own CPU
path: launch_dtw_batch_cpu / launch_frechet_batch_cpu / launch_lcss_batch_cpu
(dtw.cpp, frechet.cpp, lcss.cpp) compiled from the reference sources into
oracle/_ref/libref.so by oracle/Makefile. Run here, where the reference
sources exist:

    make -C oracle && python tests/golden/make_golden.py

Each fixture holds the inputs (t1 queries, t2 database, fixed lengths n and m,
Q == N as the reference requires) and the reference's h_results.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests.helpers import planted, walks  # noqa: E402

CASES = [
    # measure, num_t, n, m, eps, seed
    (oracle.DTW, 24, 40, 56, 0.5, 1),
    (oracle.DTW, 16, 33, 33, 0.5, 2),
    (oracle.FRECHET, 24, 56, 40, 0.5, 3),
    (oracle.FRECHET, 10, 17, 70, 0.5, 4),
    (oracle.LCSS, 24, 48, 48, 8e-4, 5),
    (oracle.LCSS, 12, 30, 61, 6e-4, 6),
]
NAMES = {oracle.DTW: "dtw", oracle.FRECHET: "frechet", oracle.LCSS: "lcss"}


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: build it with make -C oracle")
    for measure, num_t, n, m, eps, seed in CASES:
        t1, t1_off = walks(1000 + seed, [n] * num_t)
        t2, _ = planted(t1, t1_off, np.arange(num_t)[::-1], [m] * num_t, 1e-3, seed)
        res, _ = oracle.ref_batch(measure, t1, t2, num_t, n, m, eps)
        name = "ref_%s_%dx%dx%d.npz" % (NAMES[measure], num_t, n, m)
        np.savez_compressed(os.path.join(HERE, name), t1=t1, t2=t2, num_t=num_t, n=n, m=m, measure=measure,
                            eps=np.float32(eps), h_results=res)
        print(name, res[:4])


if __name__ == "__main__":
    main()
