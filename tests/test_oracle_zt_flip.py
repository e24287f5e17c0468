"""Pins of the integer-ring tryFlip (scheme_integer.cu:907-975) in the oracle.

The reference visits every candidate pair in a uniformly random order
(Fisher-Yates, random.cuh:9-22), draws two orientation coins per visited
candidate and applies the first candidate that is admissible; it fails only
when no candidate is.  Admissibility does not depend on the coins: coin bit 0
swaps j and k, which exchanges the two tests of :956-971, and bit 1 swaps the
indices, to which limitSum and the range test of limitSub are symmetric.  The
coins only choose which flip an admissible candidate applies.  So the law of
the applied flip is: a uniform pick among the admissible candidates, each of
its 4 coin settings with probability 1/4, and failure exactly when no
candidate is admissible.

The checks restate admissibility independently (integer vectors, not the
oracle's bit formulas) and compare the oracle against that law: a chi-square
test of the applied flips, and failure only for schemes without admissible
candidates -- including a scheme that hides its only admissible candidate
among 63 inadmissible ones, where a walk that gave up after 16 visits would
fail three times in four.  Both of the oracle's orders are checked: the
canonical one (the GPU's rule: one-at-a-time visits, then a reservoir pass
over the rest) and the reference's stored-list permutation.
"""
import ctypes
import random

import numpy as np
import pytest
from scipy.stats import chisquare

import oracle_ffi as O
import schemes_np as S
from minimizer_util import record_blob

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def try_flip(blob, seed, walker=0, it=0, mode=0, visit_cap=16, check_reduce=1):
    L = O.lib()
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.fgo_blob_try_flip_zt.argtypes = [i32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, i32p, i32p]
    b = np.ascontiguousarray(blob, dtype=np.int32)
    out = np.zeros(4 + 3 * 350 * 64, dtype=np.int32)
    op = np.zeros(5, dtype=np.int32)
    ok = L.fgo_blob_try_flip_zt(b.ctypes.data_as(i32p), seed, walker, it, mode, visit_cap, check_reduce,
                                out.ctypes.data_as(i32p), op.ctypes.data_as(i32p))
    assert ok >= 0
    n1, n2, n3, m = (int(x) for x in out[:4])
    return ok, tuple(int(x) for x in op), out[:4 + m * (n1 * n2 + n2 * n3 + n3 * n1)].copy()


# ---- independent restatement of the admissibility tests ----------------
def first_nonzero_positive(x):
    nz = x[x != 0]
    return len(nz) == 0 or nz[0] > 0


def limit_sum(a, b, fp):  # Addition::limitSum (addition.cu:172-186)
    s = a + b
    return bool(np.all(np.abs(s) <= 1)) and (not fp or first_nonzero_positive(s))


def limit_sub(a, b):  # Addition::limitSub (addition.cu:188-201), firstPositiveNonZero = false
    return bool(np.all(np.abs(a - b) <= 1))


def pos_sub(a, b):  # positiveFirstNonZeroSub (addition.cu:203-208)
    return first_nonzero_positive(a - b)


def candidate_ops(blob):
    """For every candidate (c, t1 < t2, equal component-c rows) in fresh-list
    order, the flip each of its 4 coin settings applies (None: inadmissible).
    Coin bit 0 swaps j and k, bit 1 swaps the two indices (scheme_integer.cu:944-954)."""
    _, U, V, W = S.from_blob(blob)
    F = (U, V, W)
    out = []
    for c, a, b in S.equal_pairs(blob):
        ops = []
        for coin in range(4):
            j, k = [x for x in range(3) if x != c]
            if coin & 1:
                j, k = k, j
            t1, t2 = (b, a) if coin & 2 else (a, b)
            op = None
            if limit_sum(F[j][t1], F[j][t2], j != 2) and limit_sub(F[k][t2], F[k][t1]):
                fwd = k == 2 or pos_sub(F[k][t2], F[k][t1])
                op = (c, j, k, t1, t2) if fwd else (c, j, k, t2, t1)
            elif limit_sum(F[k][t1], F[k][t2], k != 2) and limit_sub(F[j][t2], F[j][t1]):
                fwd = j == 2 or pos_sub(F[j][t2], F[j][t1])
                op = (c, k, j, t1, t2) if fwd else (c, k, j, t2, t1)
            ops.append(op)
        out.append(ops)
    return out


def flip_law(cands):
    """Exact probability of each applied flip and of failure."""
    a = [sum(op is not None for op in ops) / 4.0 for ops in cands]
    law = {}
    for p, ops in enumerate(cands):
        # distribution of the number of other candidates admissible under their coins
        dist = np.zeros(len(cands) + 1)
        dist[0] = 1.0
        for q, aq in enumerate(a):
            if q != p:
                dist = dist * (1 - aq) + np.concatenate([[0.0], dist[:-1]]) * aq
        e = float(np.sum(dist / (1.0 + np.arange(len(dist)))))
        for op in ops:
            if op is not None:
                law[op] = law.get(op, 0.0) + 0.25 * e
    law[None] = float(np.prod([1 - x for x in a]))
    return law


def walked_zt(n, walkers, iterations, seed):
    recs = O.naive_records(*n, walkers, ring=O.RING_ZT)
    recs = O.walk_records(recs, O.make_params(iterations=iterations, plus_iterations=400), seed, ring=O.RING_ZT,
                          nthreads=8)
    return [record_blob(r, O.RING_ZT) for r in recs]


def law_schemes():
    """Two walked 3x3x3 integer schemes with inadmissible candidates among
    their admissible ones, and a synthetic term set whose 3 admissible
    candidates sit among 63 inadmissible ones (the one-at-a-time visits
    usually miss them, so the reservoir pass decides)."""
    out = []
    for blob in walked_zt((3, 3, 3), 64, 3000, seed=77):
        cands = candidate_ops(blob)
        if 0 < sum(all(op is None for op in ops) for ops in cands) and len(cands) <= 70:
            out.append(blob)
    return out[:2] + [blocked_scheme(3)]


def test_walk_keeps_signs_and_brent_every_step():
    """Integer walks step by step: after every iteration U and V rows are
    first-nonzero-positive, every coefficient is in {-1, 0, 1} and the scheme
    is exact over Z (numpy Brent check)."""
    for n, walkers, steps in (((3, 3, 3), 16, 150), ((4, 4, 4), 8, 60)):
        recs = O.naive_records(*n, walkers, ring=O.RING_ZT)
        p = O.make_params(iterations=1, plus_iterations=40, p_expand=0.0)
        for _ in range(steps):
            recs = O.walk_records(recs, p, seed=5, ring=O.RING_ZT, nthreads=8)
            for r in recs:
                blob = record_blob(r, O.RING_ZT)
                nn, U, V, W = S.from_blob(blob)
                assert np.abs(U).max() <= 1 and np.abs(V).max() <= 1 and np.abs(W).max() <= 1
                for M in (U, V):
                    assert all(first_nonzero_positive(row) for row in M)
                assert S.brent_ok(nn, U, V, W, "ZT")
        assert int(recs["total_flips"].sum()) > 0


def test_applied_flip_is_admissible_and_keeps_brent():
    """Every flip the oracle applies is one the restated tests admit for some
    coin setting of that candidate, and the result is exact over Z."""
    blobs = walked_zt((3, 3, 3), 6, 3000, seed=78) + walked_zt((4, 4, 4), 4, 3000, seed=79)
    for q, blob in enumerate(blobs):
        allowed = {op for ops in candidate_ops(blob) for op in ops if op is not None}
        for seed in range(40):
            ok, op, out = try_flip(blob, seed=1000 + seed, walker=q)
            assert ok == 1 and op in allowed
            nn, U, V, W = S.from_blob(out)
            assert S.brent_ok(nn, U, V, W, "ZT")
            for M in (U, V):
                assert all(first_nonzero_positive(row) for row in M)


@pytest.mark.parametrize("mode,visit_cap", [(0, 16), (0, 1), (1, 16)])
def test_flip_law_matches_first_admissible_of_a_random_order(mode, visit_cap):
    """Chi-square test of the applied flips against the exact law: canonical
    order with the full visit phase, canonical with one visit (the reservoir
    pass decides whenever the first visit fails), and the reference's
    permutation order."""
    blobs = law_schemes()
    assert len(blobs) == 3
    for q, blob in enumerate(blobs):
        law = flip_law(candidate_ops(blob))
        N = 3000
        counts = {}
        for s in range(N):
            ok, op, _ = try_flip(blob, seed=7, walker=q * N + s, mode=mode, visit_cap=visit_cap)
            key = op if ok else None
            counts[key] = counts.get(key, 0) + 1
        assert set(counts) <= set(law)
        keys = [k for k in law if law[k] > 0]
        exp = np.array([law[k] * N for k in keys])
        obs = np.array([counts.get(k, 0) for k in keys])
        # pool the rare outcomes so every expected count is at least 5
        big = exp >= 5
        f_obs = np.concatenate([obs[big], [obs[~big].sum()]]) if (~big).any() else obs
        f_exp = np.concatenate([exp[big], [exp[~big].sum()]]) if (~big).any() else exp
        assert chisquare(f_obs, f_exp).pvalue > 1e-4


def test_fails_exactly_when_no_candidate_can_be_admissible():
    """Schemes without an admissible candidate make tryFlip fail for every
    seed; schemes with one never fail, however many inadmissible candidates
    surround it (more than the 16 one-at-a-time visits)."""
    seen_all = 0
    for blob in walked_zt((4, 4, 4), 32, 4000, seed=91) + walked_zt((3, 3, 3), 32, 4000, seed=92):
        cands = candidate_ops(blob)
        a = [sum(op is not None for op in ops) for ops in cands]
        if cands and max(a) == 0:
            for s in range(20):
                assert try_flip(blob, seed=s)[0] == 0
                assert try_flip(blob, seed=s, visit_cap=1)[0] == 0
        if 4 in a:
            seen_all += 1
            for s in range(20):
                assert try_flip(blob, seed=s)[0] == 1
    assert seen_all > 0
    # one admissible candidate hidden among 63 inadmissible ones
    blob = blocked_scheme(2)
    cands = candidate_ops(blob)
    a = [sum(op is not None for op in ops) for ops in cands]
    assert len(cands) > 16 and sum(x > 0 for x in a) == 1 and max(a) == 4
    for s in range(200):
        assert try_flip(blob, seed=s)[0] == 1
        assert try_flip(blob, seed=s, visit_cap=1)[0] == 1
        assert try_flip(blob, seed=s, mode=1)[0] == 1
    assert all(try_flip(blocked_scheme(0), seed=s)[0] == 0 for s in range(20))


def blocked_scheme(admissible_terms):
    """A 2x2x2 term set (not a matrix-multiplication scheme; tryFlip only needs
    the terms): 7 equal terms whose rows add or subtract to +-2, so their 63
    candidates are inadmissible in every component, and `admissible_terms`
    terms sharing the U row e1 with V and W rows e1, e2, ... that flip freely."""
    rows_u = [[0, 1, 0, 0]] * 7
    rows_v = [[1, 1, 1, 1]] * 7
    rows_w = [[1, 1, 1, 1]] * 7
    for q in range(admissible_terms):
        e = [0, 0, 0, 0]
        e[q] = 1
        rows_u = rows_u + [[1, 0, 0, 0]]
        rows_v = rows_v + [e]
        rows_w = rows_w + [e]
    return S.to_blob((2, 2, 2), rows_u, rows_v, rows_w)
