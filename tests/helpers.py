"""Shared test data: seeded random walks (the library's counter-based
generator) and planted near-duplicates (noisy, index-resampled copies of DB
trajectories, the idea of the reference's generate_similar_trajectories_nm,
main.cu:34-64, with its own generator)."""
import numpy as np

import cudaruntime1_amd as ts


def walks(seed, lengths, sigma=1e-3):
    return ts.gen_walks(seed, np.asarray(lengths, np.int32), sigma)


def uniform_lengths(seed, count, lo, hi):
    return ts.gen_lengths(seed, count, lo, hi, False)


def planted(db_xy, db_off, picks, lengths, noise, seed):
    """Query i is DB trajectory picks[i] resampled to lengths[i] points plus
    uniform noise in [-noise, noise]."""
    rng = np.random.default_rng(seed)
    out = []
    for p, L in zip(picks, lengths):
        src = db_xy[db_off[p]:db_off[p + 1]]
        m = len(src)
        idx = np.round(np.arange(L) * (m - 1) / max(L - 1, 1)).astype(np.int64) if L > 1 else np.zeros(1, np.int64)
        q = src[np.clip(idx, 0, m - 1)] + rng.uniform(-noise, noise, size=(L, 2)).astype(np.float32)
        out.append(q.astype(np.float32))
    off = ts.offsets_of([len(q) for q in out])
    return np.concatenate(out).astype(np.float32), off


def traj(xy, off, i):
    return xy[off[i]:off[i + 1]]


def say(*a):
    """Name the case before a GPU step (flushed, so a crash still shows it)."""
    print("[case]", *a, flush=True)
