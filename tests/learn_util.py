"""Helpers of the learned-proposal tests: equal-factor pairs of walker
records and random Z2 scheme batches for the scorer."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import scorer_ref as SR  # noqa: E402


def record_pairs(rec):
    """All (c, t1 < t2) with equal factors in component c of the current scheme."""
    m = int(rec["m"])
    out = []
    for c in range(3):
        v = rec["cur_v"][c][:m]
        for a in range(m):
            for b in range(a + 1, m):
                if v[a] == v[b]:
                    out.append((c, a, b))
    return out


def random_forced(recs, rng, p_none=0.2):
    """A proposed pair per walker: a random equal-factor pair, or none."""
    forced = np.zeros((len(recs), 4), dtype=np.int32)
    forced[:, 0] = -1
    for w, r in enumerate(recs):
        pr = record_pairs(r)
        if pr and rng.random() > p_none:
            forced[w, :3] = pr[rng.integers(len(pr))]
            forced[w, 3] = 1
    return forced


def random_batch(rng, count, M, nn, m_lo=None):
    """Random factor words [count][3][M], ranks and pairs (t1 < t2) for the scorer."""
    m_lo = m_lo or max(2, M // 2)
    m = rng.integers(m_lo, M + 1, size=count).astype(np.int32)
    words = np.zeros((count, 3, M), dtype=np.uint32)
    for c in range(3):
        words[:, c, :] = rng.integers(0, 1 << nn[c], size=(count, M), dtype=np.uint64).astype(np.uint32)
    words *= (np.arange(M)[None, None, :] < m[:, None, None])
    pairs = np.zeros((count, 3), dtype=np.int32)
    for q in range(count):
        a, b = sorted(rng.choice(int(m[q]), size=2, replace=False))
        pairs[q] = (rng.integers(3), a, b)
    y = rng.integers(0, 2, size=count).astype(np.float32)
    return words, m, pairs, y
