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


def z2_blob(n, U, V, W):
    return np.abs(S.to_blob(n, U, V, W)).astype(np.int32)


def schemes():
    yield "strassen", z2_blob(*S.strassen())
    yield "naive234", O.naive_blob(2, 3, 4)
    yield "walked333", walked_schemes((3, 3, 3), 8, 3000, 5, RING)[0]


def fmt(blob):
    return tuple(int(x) for x in blob[:3]), int(blob[3])


@pytest.mark.parametrize("name,blob", list(schemes()), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_operators_keep_schemes_valid(name, blob):
    assert O.validate(blob, RING) == 1
    n, m = fmt(blob)
    for p in range(3):
        for q in range(n[p]):
            out = O.resize_op(blob, "project", p, q)
            nn, mm = fmt(out)
            assert nn == tuple(n[c] - (c == p) for c in range(3)) and mm <= m
            assert O.validate(out, RING) == 1, f"project({p},{q})"
        out = O.resize_op(blob, "extend", p)
        nn, mm = fmt(out)
        j, k = (p + 1) % 3, (p + 2) % 3
        assert nn == tuple(n[c] + (c == p) for c in range(3)) and mm == m + n[j] * n[k]
        assert O.validate(out, RING) == 1, f"extend({p})"
        out = O.resize_op(blob, "product", p)
        nn, mm = fmt(out)
        assert nn == tuple(n[c] * (2 if c == p else 1) for c in range(3)) and mm == 2 * m
        assert O.validate(out, RING) == 1, f"product({p})"
        out = O.resize_op(blob, "merge", p, other=blob)
        assert fmt(out) == (tuple(n[c] * (2 if c == p else 1) for c in range(3)), 2 * m)
        assert O.validate(out, RING) == 1, f"merge({p})"
    for p1 in range(3):
        for p2 in range(3):
            out = O.resize_op(blob, "swap_size", p1, p2)
            assert O.validate(out, RING) == 1, f"swap_size({p1},{p2})"
            assert sorted(fmt(out)[0]) == sorted(n) and fmt(out)[1] == m
            assert np.array_equal(O.resize_op(out, "swap_size", p1, p2), blob)
    for i1 in range(n[0]):
        for i2 in range(n[0]):
            out = O.resize_op(blob, "swap_rows", i1, i2)
            assert O.validate(out, RING) == 1 and np.array_equal(O.resize_op(out, "swap_rows", i1, i2), blob)
    for j1 in range(n[2]):
        for j2 in range(n[2]):
            out = O.resize_op(blob, "swap_columns", j1, j2)
            assert O.validate(out, RING) == 1 and np.array_equal(O.resize_op(out, "swap_columns", j1, j2), blob)


def test_oracle_product_of_schemes_is_the_kronecker_square():
    s = S.strassen()
    out = O.resize_op(z2_blob(*s), "product_scheme", other=z2_blob(*s))
    assert np.array_equal(out, z2_blob(*S.kron_scheme(s, s)))
    assert O.validate(out, RING) == 1
    # a merge of two different formats along the free dimension
    a, b = O.naive_blob(2, 3, 2), O.naive_blob(3, 3, 2)
    out = O.resize_op(a, "merge", 0, other=b)
    assert fmt(out) == ((5, 3, 2), 12 + 18) and O.validate(out, RING) == 1


def _walked_records(count, seed, launches=2):
    recs = O.naive_records(3, 3, 3, count, ring=RING)
    return O.walk_records(recs, O.make_params(iterations=400, plus_iterations=300), seed, ring=RING,
                          launches=launches, nthreads=8)


FIELDS = ["n", "m", "best_m", "best_n", "it", "aux", "aux_idx", "total_flips", "sum_m", "flips_count"]


def same_records(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs at {np.nonzero(a[k] != b[k])[0][:5]}"
    for w in range(len(a)):
        m, bm = int(a[w]["m"]), int(a[w]["best_m"])
        assert np.array_equal(a[w]["cur_v"][:, :m], b[w]["cur_v"][:, :m]), f"{what}: walker {w} scheme"
        assert np.array_equal(a[w]["best_v"][:, :bm], b[w]["best_v"][:, :bm]), f"{what}: walker {w} best"


@pytest.mark.parametrize("prob", [1.0, 0.3])
def test_emulated_resize_matches_oracle(prob):
    seed, cap = 21, 64
    thr = (1 << 32) if prob >= 1 else int(np.floor(prob * 4294967296.0))
    recs = _walked_records(96, seed)
    emu, err = E.walk(recs, O.make_params(iterations=1), seed, resize=(thr, cap, 32))
    assert err == 0
    ora = O.resize_records(recs, seed, thr, cap)
    same_records(emu, ora, "resize")
    changed = sum(tuple(r["n"]) != (3, 3, 3) or int(r["m"]) != int(q["m"]) for r, q in zip(ora, recs))
    assert changed > (50 if prob >= 1 else 10)
    for w in range(len(ora)):
        assert O.record_valid(ora, w, ring=RING) == 1, f"walker {w} {tuple(ora[w]['n'])}"
        assert int(ora[w]["m"]) <= cap
    # the resized pool walks on, formats mixed within a block
    p = O.make_params(iterations=300, plus_iterations=300)
    e2, err = E.walk(emu, p, seed, wide=True)
    assert err == 0
    same_records(e2, O.walk_records(ora, p, seed, ring=RING), "walk after resize")


@pytest.mark.gpu
def test_gpu_resize_bit_identical_to_oracle():
    import flipgraphgpu_amd as fg
    seed, cap, W = 17, 64, 512
    print(f"[resize] 3x3x3 walkers={W} capacity={cap}", flush=True)
    e = fg.Engine(W, seed, max_iterations=400, plus_iterations=300, max_rank=cap, word_bits=32)
    try:
        e.init_naive(3, 3, 3)
        for _ in range(2):
            assert e.walk(400) == 0
        for q, prob in enumerate([1.0, 0.3, 1.0]):
            before = e.read_walkers()
            thr = (1 << 32) if prob >= 1 else int(np.floor(prob * 4294967296.0))
            assert e.resize(prob) == 0
            got = e.read_walkers()
            want = O.resize_records(before, seed, thr, cap)
            same_records(got, want, f"resize {q}")
            assert e.walk(300) == 0
            same_records(e.read_walkers(), O.walk_records(want, O.make_params(iterations=300, plus_iterations=300),
                                                          seed, ring=RING, nthreads=8), f"walk {q}")
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


@pytest.mark.parametrize("ps,pb", [(20.0, 0.0), (0.0, 20.0), (10.0, 10.0)])
def test_sandwiching_and_basis_walk_emulation_matches_oracle(ps, pb):
    # sandwiching and swapBasis (src/schemes/scheme_z2.cu:1277-1313) keep every
    # scheme valid: a consistent change of basis of the three matrices
    seed = 8
    recs = _walked_records(64, seed, launches=1)
    p = _meta_params(ps, pb)
    ora = O.walk_records(recs, p, seed, ring=RING, nthreads=8)
    emu, err = E.walk(recs, p, seed, wide=True)
    assert err == 0
    same_records(emu, ora, f"meta walk {ps} {pb}")
    for w in range(len(ora)):
        assert O.record_valid(ora, w, ring=RING) == 1 and O.record_valid(ora, w, best=True, ring=RING) == 1
    # the edges did fire: the same walk without them diverges
    plain = O.walk_records(recs, O.make_params(iterations=300, plus_iterations=300), seed, ring=RING, nthreads=8)
    assert not np.array_equal(plain["aux"], ora["aux"])


@pytest.mark.gpu
def test_gpu_sandwiching_and_basis_walk_bit_identical():
    import flipgraphgpu_amd as fg
    seed, W = 8, 256
    print("[resize] meta walk 3x3x3", flush=True)
    e = fg.Engine(W, seed, max_iterations=300, plus_iterations=300, sandwiching_probability=10.0,
                  basis_probability=10.0)
    try:
        e.init_naive(3, 3, 3)
        recs = e.read_walkers()
        for q in range(2):
            assert e.walk(300) == 0
            recs = O.walk_records(recs, _meta_params(10.0, 10.0), seed, ring=RING, nthreads=8)
            same_records(e.read_walkers(), recs, f"meta launch {q}")
    finally:
        e.close()
