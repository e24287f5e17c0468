"""The library's host Brent check and complexity (fg_verify_host,
fg_complexity_host; no GPU needed) against the independent numpy Brent
check and the closed-form complexities, and against the oracle on walked
schemes with single-coefficient corruptions."""
import numpy as np

import flipgraphgpu_amd as fg
import oracle_ffi as O
import schemes_np as S
from minimizer_util import record_blob


def test_textbook_schemes():
    for ring in (fg.RING_Z2, fg.RING_ZT):
        blobs = [S.to_blob(*S.strassen()), S.to_blob(*S.strassen_squared()), O.naive_blob(3, 3, 3, ring=ring)]
        assert fg.verify_host(blobs, ring).all()
    # naive n1 x n2 x n3: n1 n3 (n2 - 1) additions; Strassen: 18
    cx = fg.complexity_host([O.naive_blob(2, 3, 4), O.naive_blob(4, 4, 4), S.to_blob(*S.strassen())])
    assert cx.tolist() == [2 * 4 * 2, 4 * 4 * 3, 18]


def test_walked_and_corrupted_schemes_match_numpy_and_oracle():
    rng = np.random.default_rng(5)
    for ring in (O.RING_Z2, O.RING_ZT):
        recs = O.walk_records(O.naive_records(3, 3, 3, 24, ring=ring), O.make_params(iterations=800), 9, ring=ring,
                              nthreads=8)
        blobs = []
        for r in recs:
            b = record_blob(r, ring)
            blobs.append(b)
            c = b.copy()
            q = int(rng.integers(4, len(c)))
            c[q] = (c[q] ^ 1) if ring == O.RING_Z2 else {0: 1, 1: -1, -1: 0}[int(c[q])]
            blobs.append(c)
        ok = fg.verify_host(blobs, ring)
        for b, k in zip(blobs, ok):
            n, U, V, W = S.from_blob(b)
            assert bool(k) == S.brent_ok(n, U, V, W, "Z2" if ring == O.RING_Z2 else "ZT")
            assert bool(k) == (O.validate(b, ring) == 1)
        cx = fg.complexity_host(blobs, ring)
        assert cx.tolist() == [O.complexity(b, ring) for b in blobs]
        assert 0 < ok.sum() < len(blobs)
