"""Walk-level pins for the oracle: the history-ordered FlipSet (reference
mode) always equals the brute-force pair set, every state stays Brent-valid,
and canonical and reference pair orders give the same rank statistics."""
import ctypes

import numpy as np
import pytest

import oracle_ffi as O


def _check(ring, n, iters, walkers, plus):
    L = O.lib()
    L.fgo_check_flipset_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    return L.fgo_check_flipset_walk(ring, *n, 17, iters, walkers, plus)


@pytest.mark.parametrize("n", [(2, 2, 2), (2, 2, 3), (3, 3, 3)])
def test_reference_flipset_tracks_true_pair_set_z2(n):
    assert _check(O.RING_Z2, n, 600, 6, 50) == 0


@pytest.mark.parametrize("n", [(2, 2, 2), (2, 3, 3)])
def test_reference_flipset_tracks_true_pair_set_zt(n):
    assert _check(O.RING_ZT, n, 300, 4, 40) == 0


def test_walk_222_reaches_rank7_all_valid():
    recs = O.naive_records(2, 2, 2, 128)
    recs = O.walk_records(recs, O.make_params(iterations=3000, plus_iterations=10000), seed=3)
    assert (recs["best_m"] == 7).mean() > 0.8
    for i in range(len(recs)):
        assert O.record_valid(recs, i) == 1
        assert O.record_valid(recs, i, best=True) == 1
    assert (recs["it"] == 3000).all()


def test_walk_is_deterministic_and_resumable():
    p = O.make_params(iterations=700, plus_iterations=300, p_expand=0.5)
    a = O.walk_records(O.naive_records(3, 3, 3, 16), p, seed=9, launches=2)
    b = O.naive_records(3, 3, 3, 16)
    b = O.walk_records(b, p, seed=9)
    b = O.walk_records(b, p, seed=9)
    assert a.tobytes() == b.tobytes()


def test_canonical_and_reference_order_same_statistics():
    # same pair set, same uniform draw: the rank distribution after a fixed
    # budget must agree within sampling error
    p_c = O.make_params(iterations=4000, plus_iterations=2000, select_mode=O.SELECT_CANONICAL)
    p_r = O.make_params(iterations=4000, plus_iterations=2000, select_mode=O.SELECT_REFERENCE)
    a = O.walk_records(O.naive_records(3, 3, 3, 384), p_c, seed=21, nthreads=8)
    b = O.walk_records(O.naive_records(3, 3, 3, 384), p_r, seed=22, nthreads=8)
    ma, mb = a["best_m"].astype(float), b["best_m"].astype(float)
    se = np.sqrt(ma.var() / len(ma) + mb.var() / len(mb))
    assert abs(ma.mean() - mb.mean()) < 4 * se + 0.05, (ma.mean(), mb.mean(), se)


def test_zt_walk_valid_over_integers():
    p = O.make_params(iterations=1500, plus_iterations=500, p_expand=0.2)
    recs = O.walk_records(O.naive_records(2, 2, 2, 32, ring=O.RING_ZT), p, seed=5, ring=O.RING_ZT)
    for i in range(len(recs)):
        assert O.record_valid(recs, i, ring=O.RING_ZT) == 1
        assert O.record_valid(recs, i, best=True, ring=O.RING_ZT) == 1
    assert recs["best_m"].min() <= 7
