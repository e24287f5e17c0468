"""Complexity minimizer (src/entities/complexity_minimizer.cu) on the CPU:
the oracle's minimize mode pinned by the additions count of known schemes
and by its step semantics, and the minimizer kernel's own lane code, run on
the host, matched bit-for-bit against the oracle."""
import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
import schemes_np as S
from minimizer_util import broadcast, minimizer_start, record_complexity
from walk_compare import first_mismatch


@pytest.mark.parametrize("n", [(2, 2, 2), (2, 3, 4), (3, 3, 3), (4, 4, 4), (3, 4, 5)])
def test_complexity_of_naive_is_its_addition_count(n):
    # every entry of C = A B sums n2 products: n2 - 1 additions each
    n1, n2, n3 = n
    for ring in (O.RING_Z2, O.RING_ZT):
        assert O.complexity(O.naive_blob(*n, ring=ring), ring) == n1 * n3 * (n2 - 1)


def test_complexity_of_strassen_is_18():
    blob = S.to_blob(*S.strassen())
    assert O.complexity(blob, O.RING_ZT) == 18
    assert O.complexity(blob, O.RING_Z2) == 18


def _walked(n, walkers, iterations, seed, ring):
    """Walkers a few thousand flips away from the naive scheme (lower rank, mixed complexity)."""
    recs = O.naive_records(*n, walkers, ring=ring)
    p = O.make_params(iterations=iterations, plus_iterations=400)
    return O.walk_records(recs, p, seed, ring=ring, nthreads=8)


@pytest.mark.parametrize("ring", [O.RING_Z2, O.RING_ZT])
def test_oracle_minimize_steps(ring):
    start, cx0 = minimizer_start(_walked((3, 3, 3), 16, 1500, 5, ring), ring)
    p1 = O.make_params(iterations=1, mode=1, check_reduce=False)
    steps = 300
    cur = start
    improved = 0
    for _ in range(steps):
        nxt = O.walk_records(cur, p1, 77, ring=ring)
        for w in range(len(cur)):
            flipped = nxt[w]["total_flips"] - cur[w]["total_flips"]
            assert nxt[w]["it"] == cur[w]["it"] + 1
            if not flipped:
                # a failed flip changes nothing
                assert nxt[w]["m"] == cur[w]["m"] and np.array_equal(nxt[w]["cur_v"], cur[w]["cur_v"])
                continue
            assert nxt[w]["sum_m"] - cur[w]["sum_m"] == cur[w]["m"]
            assert nxt[w]["m"] <= cur[w]["m"]  # no expansions; only zero factors shrink it
            cx = record_complexity(nxt[w], ring)
            if cx < cur[w]["best_rank"]:
                improved += 1
                assert nxt[w]["best_rank"] == cx
                assert nxt[w]["best_m"] == nxt[w]["m"]
                m = int(nxt[w]["m"])
                assert np.array_equal(nxt[w]["best_v"][:, :m], nxt[w]["cur_v"][:, :m])
            else:
                assert nxt[w]["best_rank"] == cur[w]["best_rank"]
                assert np.array_equal(nxt[w]["best_v"], cur[w]["best_v"])
        cur = nxt
    assert improved > 0
    assert (cur["best_rank"] <= cx0).all()
    for w in range(len(cur)):
        assert O.record_valid(cur, w, best=False, ring=ring)
        assert O.record_valid(cur, w, best=True, ring=ring)
    # one launch of `steps` iterations is the same walk
    whole = O.walk_records(start, O.make_params(iterations=steps, mode=1, check_reduce=False), 77, ring=ring)
    assert first_mismatch(whole, cur) is None


def test_oracle_minimize_stops_without_pairs():
    # Strassen's scheme has no two equal factors in any component: the first
    # tryFlip fails and the minimizer stops after one draw
    blob = S.to_blob(*S.strassen())
    assert O.pairs(blob) == []
    recs = O.naive_records(2, 2, 2, 4)
    _, comps = O.blob_to_terms(blob)
    for r in recs:
        r["m"] = 7
        r["cur_v"][:] = 0
        for c in range(3):
            for t in range(7):
                r["cur_v"][c][t] = sum((abs(x) & 1) << q for q, x in enumerate(comps[c][t]))
    start, cx0 = minimizer_start(recs)
    assert cx0 == 18
    out = O.walk_records(start, O.make_params(iterations=1000, mode=1, check_reduce=False), 3)
    assert (out["it"] == 1).all() and (out["total_flips"] == 0).all()
    assert (out["best_rank"] == 18).all()


@pytest.mark.parametrize("ring", [O.RING_Z2, O.RING_ZT])
@pytest.mark.parametrize("n,walkers,iterations", [((2, 2, 3), 64, 1500), ((3, 3, 3), 64, 2000), ((4, 4, 4), 32, 1500)])
def test_emulated_minimizer_matches_oracle(ring, n, walkers, iterations):
    start, _ = minimizer_start(_walked(n, walkers, 1500, 9, ring), ring)
    p = O.make_params(iterations=iterations, mode=1, check_reduce=False)
    emu, err = E.walk(start, p, 31, ring=ring)
    ora = O.walk_records(start, p, 31, ring=ring, nthreads=8)
    assert err == 0
    assert first_mismatch(emu, ora) is None, first_mismatch(emu, ora)
    # second launch after a re-broadcast of the best walker
    src = int(np.argmin(ora["best_rank"]))
    dst = list(range(0, walkers, 3))
    emu2, err = E.walk(broadcast(emu, src, dst), p, 31, ring=ring)
    ora2 = O.walk_records(broadcast(ora, src, dst), p, 31, ring=ring, nthreads=8)
    assert err == 0
    assert first_mismatch(emu2, ora2) is None, first_mismatch(emu2, ora2)
    assert (ora2["total_flips"] > ora["total_flips"]).any()
