"""Pins the oracle's scheme operators (naive init, Brent validation,
complexity, pair order, flip / reduce / plus / split) with checks that do not
reuse the oracle's code: closed forms, a numpy restatement of the
matrix-multiplication tensor, textbook schemes and brute-force invariants."""
import itertools
import random

import numpy as np
import pytest

import oracle_ffi as O
import schemes_np as S

FORMATS = [(2, 2, 2), (2, 2, 3), (2, 3, 4), (3, 3, 3), (3, 3, 4), (3, 4, 5), (4, 4, 4), (2, 3, 2), (5, 5, 5)]


@pytest.mark.parametrize("n", FORMATS)
@pytest.mark.parametrize("ring", [O.RING_Z2, O.RING_ZT])
def test_naive_is_valid_and_matches_numpy_tensor(n, ring):
    blob = O.naive_blob(*n, ring=ring)
    nn, U, V, W = S.from_blob(blob)
    assert nn == n and U.shape[0] == n[0] * n[1] * n[2]
    assert S.brent_ok(n, U, V, W, "Z2" if ring == O.RING_Z2 else "ZT")
    assert O.validate(blob, ring) == 1


@pytest.mark.parametrize("n", FORMATS)
def test_naive_complexity_closed_form(n):
    # every naive factor has one nonzero: 3m - 2m - n3*n1
    n1, n2, n3 = n
    assert O.complexity(O.naive_blob(*n)) == n1 * n2 * n3 - n3 * n1


@pytest.mark.parametrize("n", FORMATS)
def test_naive_pair_count_closed_form(n):
    # U value a[r][l] repeats n3 times, V value b[l][q] n1 times, W value c[r][q] n2 times
    n1, n2, n3 = n
    c2 = lambda x: x * (x - 1) // 2
    want = n1 * n2 * c2(n3) + n2 * n3 * c2(n1) + n3 * n1 * c2(n2)
    blob = O.naive_blob(*n)
    assert len(O.pairs(blob)) == want
    assert O.pairs(blob) == S.equal_pairs(blob)


def test_strassen_valid_z2_and_z():
    blob = S.to_blob(*S.strassen())
    assert S.brent_ok(*S.strassen(), ring="ZT")
    assert O.validate(blob, O.RING_ZT) == 1
    assert O.validate(blob, O.RING_Z2) == 1


def test_strassen_squared_rank49_valid():
    sch = S.strassen_squared()
    assert sch[0] == (4, 4, 4) and len(sch[1]) == 49
    assert S.brent_ok(*sch, ring="ZT")
    blob = S.to_blob(*sch)
    assert O.validate(blob, O.RING_ZT) == 1
    assert O.validate(blob, O.RING_Z2) == 1


def test_single_coefficient_corruption_fails():
    rng = random.Random(5)
    for sch in (S.strassen(), S.strassen_squared()):
        base = S.to_blob(*sch)
        for _ in range(40):
            b = base.copy()
            pos = rng.randrange(4, len(b))
            b[pos] = 1 - abs(int(b[pos]))  # 0 <-> 1 (and -1 -> 0)
            n, U, V, W = S.from_blob(b)
            assert O.validate(b, O.RING_Z2) == int(S.brent_ok(n, U, V, W, "Z2"))
            assert O.validate(b, O.RING_ZT) == int(S.brent_ok(n, U, V, W, "ZT"))


def test_validate_agrees_with_numpy_on_random_schemes():
    rng = np.random.default_rng(3)
    for n in [(2, 2, 2), (2, 2, 3), (2, 3, 2)]:
        nn = (n[0] * n[1], n[1] * n[2], n[2] * n[0])
        for _ in range(60):
            m = int(rng.integers(1, 9))
            U, V, W = (rng.integers(0, 2, size=(m, k)) for k in nn)
            blob = S.to_blob(n, U, V, W)
            assert O.validate(blob, O.RING_Z2) == int(S.brent_ok(n, U, V, W, "Z2"))


def test_fresh_flipset_order_is_canonical():
    # the stored FlipSet of a freshly read scheme is in component/index1/index2 order
    for sch in (S.strassen_squared(),):
        blob = S.to_blob(*sch)
        assert O.pairs(blob, mode=1) == O.pairs(blob, mode=0) == S.equal_pairs(blob)
    for n in FORMATS[:7]:
        blob = O.naive_blob(*n)
        assert O.pairs(blob, mode=1) == S.equal_pairs(blob)


def _all_directed_flips(blob):
    for (c, a, b) in S.equal_pairs(blob):
        others = [x for x in range(3) if x != c]
        for j, k in (others, others[::-1]):
            for i1, i2 in ((a, b), (b, a)):
                yield (c, j, k, i1, i2)


def test_all_48_flips_of_naive_222_preserve_brent_and_invert():
    blob = O.naive_blob(2, 2, 2)
    flips = list(_all_directed_flips(blob))
    assert len(flips) == 48
    for (i, j, k, i1, i2) in flips:
        out = O.replay(blob, [(0, i, j, k, i1, i2, 1)])
        n, U, V, W = S.from_blob(out)
        assert S.brent_ok(n, U, V, W, "Z2")
        if int(out[3]) == int(blob[3]):
            back = O.replay(out, [(0, i, j, k, i1, i2, 1)])
            assert np.array_equal(back, blob)


def test_flip_update_rule_matches_definition():
    # Z2 flip: v_j[i1] += v_j[i2]; v_k[i2] += v_k[i1] (checkReduce off, so no merge)
    rng = random.Random(1)
    blob = O.naive_blob(3, 3, 3)
    checked = 0
    for step in range(120):
        prs = S.equal_pairs(blob)
        c, a, b = rng.choice(prs)
        j, k = [x for x in range(3) if x != c]
        if rng.random() < 0.5:
            j, k = k, j
        n, U, V, W = S.from_blob(blob)
        mats = [U, V, W]
        out = O.replay(blob, [(0, c, j, k, a, b, 0)])
        want = [M.copy() for M in mats]
        want[j][a] = (want[j][a] + want[j][b]) % 2
        want[k][b] = (want[k][b] + want[k][a]) % 2
        if np.any(want[j][a]) and np.any(want[k][b]):
            got = S.from_blob(out)
            assert all(np.array_equal(got[1 + q], want[q]) for q in range(3))
            checked += 1
            blob = out
    assert checked > 50


def test_reduction_merges_terms_and_keeps_brent():
    # Two naive 2x2x2 terms share U; make them share V too by flipping, the
    # oracle must then merge (rank drops by exactly one) and stay valid.
    rng = random.Random(7)
    seen = 0
    for trial in range(60):
        blob = O.naive_blob(2, 2, 2)
        for _ in range(80):
            prs = S.equal_pairs(blob)
            if not prs:
                break
            c, a, b = rng.choice(prs)
            j, k = [x for x in range(3) if x != c]
            if rng.random() < 0.5:
                j, k = k, j
            if rng.random() < 0.5:
                a, b = b, a
            before = int(blob[3])
            blob = O.replay(blob, [(0, c, j, k, a, b, 1)])
            n, U, V, W = S.from_blob(blob)
            assert S.brent_ok(n, U, V, W, "Z2")
            assert int(blob[3]) in (before, before - 1, before - 2)
            if int(blob[3]) == before - 1:
                seen += 1
    assert seen >= 3


def test_plus_raises_rank_by_one_and_keeps_brent():
    blob = S.to_blob(*S.strassen_squared())
    rng = random.Random(11)
    for ring in (O.RING_Z2, O.RING_ZT):
        for _ in range(30):
            m = int(blob[3])
            while True:
                a, b = rng.randrange(m), rng.randrange(m)
                n, U, V, W = S.from_blob(blob)
                if a != b and not any(np.array_equal(M[a], M[b]) for M in (U, V, W)):
                    break
            i, j, k = rng.sample(range(3), 3)
            out = O.replay(blob, [(1, i, j, k, a, b, rng.randrange(3))], ring=ring)
            n, U, V, W = S.from_blob(out)
            assert S.brent_ok(n, U, V, W, "Z2" if ring == O.RING_Z2 else "ZT")
            if ring == O.RING_Z2:
                assert int(out[3]) == m + 1


def test_split_keeps_brent():
    blob = O.naive_blob(3, 3, 3)
    rng = random.Random(2)
    for _ in range(30):
        i, j, k = rng.sample(range(3), 3)
        idx = rng.randrange(27)
        a1 = rng.randrange(1, 512)
        out = O.replay(blob, [(2, i, j, k, idx, a1, 0)])
        n, U, V, W = S.from_blob(out)
        assert int(out[3]) == 28 and S.brent_ok(n, U, V, W, "Z2")


def test_zt_flips_keep_brent_over_integers():
    """Flips admitted by an independent restatement of tryFlip's tests
    (test_oracle_zt_flip.candidate_ops), applied one at a time from the naive
    3x3x3 and 2x3x4 integer schemes, keep the scheme exact over Z at every
    step, with coefficients in {-1, 0, 1} and U, V rows first-nonzero-positive.
    (Strassen^2 has no two equal factors, so it admits no flip at all.)"""
    from test_oracle_zt_flip import candidate_ops, first_nonzero_positive

    for n in ((3, 3, 3), (2, 3, 4)):
        _flip_steps(O.naive_blob(*n, ring=O.RING_ZT), candidate_ops, first_nonzero_positive)


def _flip_steps(blob, candidate_ops, first_nonzero_positive):
    rng = random.Random(4)
    steps = 0
    for _ in range(120):
        ops = [op for c in candidate_ops(blob) for op in c if op is not None]
        if not ops:
            break
        i, j, k, a, b = rng.choice(ops)
        out = O.replay(blob, [(0, i, j, k, a, b, 1)], ring=O.RING_ZT)
        n, U, V, W = S.from_blob(out)
        assert np.abs(U).max() <= 1 and np.abs(V).max() <= 1 and np.abs(W).max() <= 1
        assert all(first_nonzero_positive(row) for M in (U, V) for row in M)
        assert S.brent_ok(n, U, V, W, "ZT")
        assert O.validate(out, O.RING_ZT) == 1
        blob = out
        steps += 1
    assert steps >= 20


def test_fix_signs_makes_u_and_v_first_nonzero_positive():
    blob = S.to_blob(*S.strassen())
    out = O.replay(blob, [(5,)], ring=O.RING_ZT)
    n, U, V, W = S.from_blob(out)
    for M in (U, V):
        for row in M:
            nzr = row[row != 0]
            assert len(nzr) == 0 or nzr[0] > 0
    assert S.brent_ok(n, U, V, W, "ZT")
