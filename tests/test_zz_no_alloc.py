"""fg_verify_noalloc (the Brent check with its staging in the library's fixed
per-GPU scratch array) against fg_verify on the same blobs, through the C
ABI: two calls in a row, the second with a smaller input, both rings, valid
schemes and single-coefficient corruptions; and the size limit."""
import ctypes

import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
import schemes_np as S

pytestmark = pytest.mark.gpu


def _call(fn, ring, blobs):
    flat = np.concatenate([np.array([len(blobs)], dtype=np.int32)] + [np.asarray(b, dtype=np.int32) for b in blobs])
    ok = np.full(len(blobs), -7, dtype=np.int32)
    rc = fn(ring, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(flat),
            ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return rc, ok


def _corrupt(blob, rng, ring):
    out = np.array(blob, dtype=np.int32, copy=True)
    q = int(rng.integers(4, len(out)))
    if ring == O.RING_Z2:
        out[q] ^= 1
    else:
        out[q] = {0: 1, 1: -1, -1: 0}[int(out[q])]
    return out


def _blobs(ring, rng, walkers, n):
    recs = O.naive_records(*n, walkers, ring=ring)
    recs = O.walk_records(recs, O.make_params(iterations=2000, plus_iterations=300), seed=3, ring=ring, nthreads=8)
    from minimizer_util import record_blob
    good = [record_blob(r, ring) for r in recs]
    bad = [_corrupt(b, rng, ring) for b in good[: walkers // 2]]
    out = good + bad
    rng.shuffle(out)
    return out


@pytest.mark.parametrize("ring", [O.RING_Z2, O.RING_ZT])
def test_noalloc_matches_fg_verify_on_two_calls(ring):
    L = fg.load_library()
    rng = np.random.default_rng(11 + ring)
    big = _blobs(ring, rng, 96, (4, 4, 4)) + _blobs(ring, rng, 32, (3, 3, 3))
    big.append(S.to_blob(*S.strassen()))
    small = _blobs(ring, rng, 8, (2, 3, 4))
    for blobs in (big, small):
        rc_ref, ok_ref = _call(L.fg_verify, ring, blobs)
        rc, ok = _call(L.fg_verify_noalloc, ring, blobs)
        assert rc_ref >= 0 and rc == rc_ref, fg.load_library().fg_last_error()
        assert np.array_equal(ok != 0, ok_ref != 0)
        assert 0 < rc < len(blobs)  # both valid and corrupted schemes are present


def test_noalloc_rejects_more_than_its_scratch():
    L = fg.load_library()
    blob = O.naive_blob(4, 4, 4)
    rc, _ = _call(L.fg_verify_noalloc, O.RING_Z2, [blob] * 900)
    assert rc < 0 and b"scratch" in L.fg_last_error()
    # the array still serves a call that fits
    rc, ok = _call(L.fg_verify_noalloc, O.RING_Z2, [blob] * 4)
    assert rc == 4 and ok.tolist() == [1, 1, 1, 1]
