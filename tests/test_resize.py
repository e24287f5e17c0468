"""Resize and basis operators of the Z2 walk (src/schemes/scheme_z2.cu:184-310,
643-979,1165-1324) and resizeKernel (src/entities/flip_graph.cu:575-613).

The oracle's operators are pinned by what the mathematics fixes: a
projection, extension, dimension product, merge, product of schemes or
transposition of a valid scheme is a valid scheme of the new format, with the
rank the construction gives; transpositions and basis swaps undo themselves;
the product of Strassen's scheme with itself is the Kronecker square built
independently in schemes_np.  The kernel's resize mode (run on the host by
the emulation, and on the GPU) must then match the oracle bit for bit."""
import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
import schemes_np as S
from minimizer_util import walked_schemes

RING = O.RING_Z2
ZT = O.RING_ZT


def z2_blob(n, U, V, W):
    return np.abs(S.to_blob(n, U, V, W)).astype(np.int32)


def zt_blob(n, U, V, W):
    return O.replay(S.to_blob(n, U, V, W), [[5]], ring=ZT)  # fixSigns, as a reader applies it


def schemes():
    yield "strassen", z2_blob(*S.strassen()), RING
    yield "naive234", O.naive_blob(2, 3, 4), RING
    yield "walked333", walked_schemes((3, 3, 3), 8, 3000, 5, RING)[0], RING
    yield "strassen_zt", zt_blob(*S.strassen()), ZT
    yield "naive223_zt", O.naive_blob(2, 2, 3, ring=ZT), ZT
    yield "walked333_zt", O.replay(walked_schemes((3, 3, 3), 8, 3000, 5, ZT)[0], [[5]], ring=ZT), ZT


def fmt(blob):
    return tuple(int(x) for x in blob[:3]), int(blob[3])


@pytest.mark.parametrize("name,blob,ring", list(schemes()), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_operators_keep_schemes_valid(name, blob, ring):
    assert O.validate(blob, ring) == 1
    n, m = fmt(blob)
    op = lambda *a, **k: O.resize_op(*a, ring=ring, **k)
    # integer schemes are compared in the sign convention of fixSigns, which a
    # reader applies (scheme_integer.cu:398-425)
    canon = (lambda x: O.replay(x, [[5]], ring=ZT)) if ring == ZT else (lambda x: x)
    for p in range(3):
        for q in range(n[p]):
            out = op(blob, "project", p, q)
            nn, mm = fmt(out)
            assert nn == tuple(n[c] - (c == p) for c in range(3)) and mm <= m
            assert O.validate(out, ring) == 1, f"project({p},{q})"
        out = op(blob, "extend", p)
        nn, mm = fmt(out)
        j, k = (p + 1) % 3, (p + 2) % 3
        assert nn == tuple(n[c] + (c == p) for c in range(3)) and mm == m + n[j] * n[k]
        assert O.validate(out, ring) == 1, f"extend({p})"
        out = op(blob, "product", p)
        nn, mm = fmt(out)
        assert nn == tuple(n[c] * (2 if c == p else 1) for c in range(3)) and mm == 2 * m
        assert O.validate(out, ring) == 1, f"product({p})"
        out = op(blob, "merge", p, other=blob)
        assert fmt(out) == (tuple(n[c] * (2 if c == p else 1) for c in range(3)), 2 * m)
        assert O.validate(out, ring) == 1, f"merge({p})"
    for p1 in range(3):
        for p2 in range(3):
            out = op(blob, "swap_size", p1, p2)
            assert O.validate(out, ring) == 1, f"swap_size({p1},{p2})"
            assert sorted(fmt(out)[0]) == sorted(n) and fmt(out)[1] == m
            assert np.array_equal(canon(op(out, "swap_size", p1, p2)), canon(blob))
    for i1 in range(n[0]):
        for i2 in range(n[0]):
            out = op(blob, "swap_rows", i1, i2)
            assert O.validate(out, ring) == 1 and np.array_equal(canon(op(out, "swap_rows", i1, i2)), canon(blob))
    for j1 in range(n[2]):
        for j2 in range(n[2]):
            out = op(blob, "swap_columns", j1, j2)
            assert O.validate(out, ring) == 1 and np.array_equal(canon(op(out, "swap_columns", j1, j2)), canon(blob))


def test_oracle_integer_product_of_schemes_is_the_signed_kronecker_square():
    f = zt_blob(*S.strassen())
    n, U, V, W = S.from_blob(f)
    out = O.resize_op(f, "product_scheme", other=f, ring=ZT)
    assert np.array_equal(out, S.to_blob(*S.kron_scheme((n, U, V, W), (n, U, V, W))))
    assert O.validate(out, ZT) == 1


def test_oracle_product_of_schemes_is_the_kronecker_square():
    s = S.strassen()
    out = O.resize_op(z2_blob(*s), "product_scheme", other=z2_blob(*s))
    assert np.array_equal(out, z2_blob(*S.kron_scheme(s, s)))
    assert O.validate(out, RING) == 1
    # a merge of two different formats along the free dimension
    a, b = O.naive_blob(2, 3, 2), O.naive_blob(3, 3, 2)
    out = O.resize_op(a, "merge", 0, other=b)
    assert fmt(out) == ((5, 3, 2), 12 + 18) and O.validate(out, RING) == 1


def _walked_records(count, seed, launches=2, ring=RING):
    recs = O.naive_records(3, 3, 3, count, ring=ring)
    return O.walk_records(recs, O.make_params(iterations=400, plus_iterations=300), seed, ring=ring,
                          launches=launches, nthreads=8)


FIELDS = ["n", "m", "best_m", "best_n", "it", "aux", "aux_idx", "total_flips", "sum_m", "flips_count"]


def same_records(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs at {np.nonzero(a[k] != b[k])[0][:5]}"
    for w in range(len(a)):
        m, bm = int(a[w]["m"]), int(a[w]["best_m"])
        for k in ("cur_v", "cur_s"):
            assert np.array_equal(a[w][k][:, :m], b[w][k][:, :m]), f"{what}: walker {w} {k}"
        for k in ("best_v", "best_s"):
            assert np.array_equal(a[w][k][:, :bm], b[w][k][:, :bm]), f"{what}: walker {w} {k}"


@pytest.mark.parametrize("ring", [RING, ZT], ids=["z2", "zt"])
@pytest.mark.parametrize("prob", [1.0, 0.3])
def test_emulated_resize_matches_oracle(prob, ring):
    seed, cap = 21, 64
    coef = 16 if ring == ZT else 32
    thr = (1 << 32) if prob >= 1 else int(np.floor(prob * 4294967296.0))
    recs = _walked_records(96, seed, ring=ring)
    emu, err = E.walk(recs, O.make_params(iterations=1), seed, ring=ring, resize=(thr, cap, coef))
    assert err == 0
    ora = O.resize_records(recs, seed, thr, cap, coef, ring=ring)
    same_records(emu, ora, "resize")
    changed = sum(tuple(r["n"]) != (3, 3, 3) or int(r["m"]) != int(q["m"]) for r, q in zip(ora, recs))
    # with factors of at most 16 coefficients (integer ring) most 3x3x3 products and merges are too wide
    assert changed > (40 if prob >= 1 and ring == RING else 5)
    for w in range(len(ora)):
        assert O.record_valid(ora, w, ring=ring) == 1, f"walker {w} {tuple(ora[w]['n'])}"
        assert int(ora[w]["m"]) <= cap
    # the resized pool walks on, formats mixed within a block
    p = O.make_params(iterations=300, plus_iterations=300)
    e2, err = E.walk(emu, p, seed, ring=ring, wide=True)
    assert err == 0
    same_records(e2, O.walk_records(ora, p, seed, ring=ring), "walk after resize")


@pytest.mark.gpu
@pytest.mark.parametrize("ring", [RING, ZT], ids=["z2", "zt"])
def test_gpu_resize_bit_identical_to_oracle(ring):
    import flipgraphgpu_amd as fg
    seed, cap, W = 17, 64, 512
    print(f"[resize] 3x3x3 ring={ring} walkers={W} capacity={cap}", flush=True)
    e = fg.Engine(W, seed, ring=ring, max_iterations=400, plus_iterations=300, max_rank=cap, word_bits=32)
    try:
        e.init_naive(3, 3, 3)
        for _ in range(2):
            assert e.walk(400) == 0
        for q, prob in enumerate([1.0, 0.3, 1.0]):
            before = e.read_walkers()
            thr = (1 << 32) if prob >= 1 else int(np.floor(prob * 4294967296.0))
            assert e.resize(prob) == 0
            got = e.read_walkers()
            want = O.resize_records(before, seed, thr, cap, ring=ring)
            same_records(got, want, f"resize {q}")
            assert e.walk(300) == 0
            same_records(e.read_walkers(), O.walk_records(want, O.make_params(iterations=300, plus_iterations=300),
                                                          seed, ring=ring, nthreads=8), f"walk {q}")
        # updateRanks' copy of current schemes into the bests
        recs = e.read_walkers()
        pick = [3, 100, 511]
        e.snapshot_best(pick)
        after = e.read_walkers()
        for w in pick:
            assert int(after[w]["best_m"]) == int(recs[w]["m"]) and tuple(after[w]["best_n"]) == tuple(recs[w]["n"])
            m = int(recs[w]["m"])
            assert np.array_equal(after[w]["best_v"][:, :m], recs[w]["cur_v"][:, :m])
        ranks = np.arange(W, dtype=np.int32) % 50 + 1
        e.set_best_ranks(ranks)
        assert np.array_equal(e.read_walkers()["best_rank"], ranks)
    finally:
        e.close()


def _meta_params(p_sandwich, p_basis):
    # per-iteration probabilities of a few percent: the edges fire often
    return O.make_params(iterations=300, max_iterations=300, plus_iterations=300, p_sandwich=p_sandwich,
                         p_basis=p_basis)


@pytest.mark.parametrize("ring", [RING, ZT], ids=["z2", "zt"])
@pytest.mark.parametrize("ps,pb", [(20.0, 0.0), (0.0, 20.0), (10.0, 10.0)])
def test_sandwiching_and_basis_walk_emulation_matches_oracle(ps, pb, ring):
    # sandwiching and swapBasis (src/schemes/scheme_z2.cu:1277-1313) keep every
    # scheme valid: a consistent change of basis of the three matrices
    seed = 8
    recs = _walked_records(64, seed, launches=1, ring=ring)
    p = _meta_params(ps, pb)
    ora = O.walk_records(recs, p, seed, ring=ring, nthreads=8)
    emu, err = E.walk(recs, p, seed, ring=ring, wide=True)
    assert err == 0
    same_records(emu, ora, f"meta walk {ps} {pb}")
    for w in range(len(ora)):
        assert O.record_valid(ora, w, ring=ring) == 1 and O.record_valid(ora, w, best=True, ring=ring) == 1
    # the edges did fire: the same walk without them diverges
    plain = O.walk_records(recs, O.make_params(iterations=300, plus_iterations=300), seed, ring=ring, nthreads=8)
    assert not np.array_equal(plain["aux"], ora["aux"])


@pytest.mark.gpu
@pytest.mark.parametrize("ring", [RING, ZT], ids=["z2", "zt"])
def test_gpu_sandwiching_and_basis_walk_bit_identical(ring):
    import flipgraphgpu_amd as fg
    seed, W = 8, 256
    print(f"[resize] meta walk 3x3x3 ring={ring}", flush=True)
    e = fg.Engine(W, seed, ring=ring, max_iterations=300, plus_iterations=300, sandwiching_probability=10.0,
                  basis_probability=10.0)
    try:
        e.init_naive(3, 3, 3)
        recs = e.read_walkers()
        for q in range(2):
            assert e.walk(300) == 0
            recs = O.walk_records(recs, _meta_params(10.0, 10.0), seed, ring=ring, nthreads=8)
            same_records(e.read_walkers(), recs, f"meta launch {q}")
    finally:
        e.close()
