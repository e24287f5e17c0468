"""GPU walk kernel vs the CPU oracle: per-walker trajectories must be
bit-identical (same Philox streams, same canonical pair order), through the
C ABI.  Also checks the GPU Brent check against the oracle's."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
import schemes_np as S
from walk_compare import first_mismatch

pytestmark = pytest.mark.gpu


def _run_both(n, walkers, iterations, seed, launches=1, plus=8000, p_expand=0.0, random_iterations=False,
              ring=O.RING_Z2, p_reduce=0.0, word_bits=0, partner_layout=0, look=(0, 0)):
    print(f"[parity] ring={ring} n={n} walkers={walkers} iterations={iterations} launches={launches} seed={seed} "
          f"plus={plus} p_expand={p_expand} p_reduce={p_reduce} random={random_iterations} "
          f"word_bits={word_bits} layout={partner_layout} look={look}", flush=True)
    with fg.Engine(walkers, seed, ring=ring, max_iterations=iterations, plus_iterations=plus,
                   expand_probability=p_expand, reduce_probability=p_reduce,
                   random_iterations=random_iterations, zt_visit_cap=16, word_bits=word_bits,
                   partner_layout=partner_layout, lookahead_period=look[0], lookahead_candidates=look[1]) as e:
        e.init_naive(*n)
        init = e.read_walkers()
        ora0 = O.naive_records(*n, walkers, ring=ring)
        assert first_mismatch(init, ora0) is None
        for _ in range(launches):
            assert e.walk(iterations) == 0
        gpu = e.read_walkers()
    p = O.make_params(iterations=iterations, plus_iterations=plus, p_expand=p_expand, p_reduce=p_reduce,
                      random_iterations=random_iterations, look_period=look[0], look_candidates=look[1])
    ora = O.walk_records(ora0, p, seed, launches=launches, nthreads=8, ring=ring)
    return gpu, ora


@pytest.mark.parametrize("n,walkers,iterations", [
    ((2, 2, 2), 128, 3000),
    ((2, 2, 3), 128, 3000),
    ((3, 3, 3), 128, 4000),
    ((2, 3, 4), 64, 3000),
    ((4, 4, 4), 64, 2000),
])
def test_walk_bit_identical_to_oracle(n, walkers, iterations):
    gpu, ora = _run_both(n, walkers, iterations, seed=12345)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
    assert (ora["total_flips"] > 0).all()


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (3, 3, 4)])
def test_walk_with_plus_and_expand_bit_identical(n):
    # small plateau and a high expand rate exercise plus / split / rebuild paths
    gpu, ora = _run_both(n, 128, 2500, seed=7, launches=2, plus=150, p_expand=200.0)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (2, 3, 4), (4, 4, 4)])
def test_walk_16bit_words_bit_identical(n):
    # Z2 factors of up to 16 coefficients held in 16-bit LDS words
    gpu, ora = _run_both(n, 128, 2000, seed=31, launches=2, plus=200, p_expand=100.0, word_bits=16)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (2, 3, 4)])
@pytest.mark.parametrize("layout,word_bits", [(1, 0), (1, 32), (2, 0), (3, 0)])
def test_walk_partner_layouts_bit_identical(n, layout, word_bits):
    # schemes of up to 32 terms: class masks (the default) and partner lists
    # with 16- or 32-bit words give the same trajectories
    gpu, ora = _run_both(n, 128, 2000, seed=57, launches=2, plus=150, p_expand=100.0, word_bits=word_bits,
                         partner_layout=layout)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


@pytest.mark.parametrize("n,layout", [((3, 3, 3), 2), ((3, 3, 3), 1), ((4, 4, 4), 0), ((3, 3, 3), 3)])
def test_walk_reduce_probability_bit_identical(n, layout):
    # --reduce-probability > 0: tryReduce after flips (flip_graph.cu:550-551)
    gpu, ora = _run_both(n, 128, 2000, seed=66, launches=2, plus=200, p_expand=50.0, p_reduce=400.0,
                         partner_layout=layout)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


@pytest.mark.parametrize("n,layout,look", [((3, 3, 3), 0, (2, 3)), ((3, 3, 3), 1, (1, 2)), ((4, 4, 4), 0, (2, 3)),
                                           ((2, 3, 4), 0, (4, 8)), ((3, 3, 3), 3, (2, 3))])
def test_walk_lookahead_bit_identical(n, layout, look):
    # reduction-seeking lookahead proposals (period, candidates)
    gpu, ora = _run_both(n, 128, 2000, seed=81, launches=2, plus=200, p_expand=50.0, partner_layout=layout, look=look)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_walk_5x5x5_bit_identical():
    # 25-coefficient factors (32-bit words), 125 terms, plus and expand
    gpu, ora = _run_both((5, 5, 5), 64, 600, seed=125, launches=2, plus=200, p_expand=200.0)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
    assert (ora["total_flips"] > 0).all()


@pytest.mark.parametrize("layout", [0, 3])
def test_random_iterations_bit_identical(layout):
    # 96 walkers are three queue blocks of 32: the one-lane layout's second
    # wave serves one block and leaves its upper half-wave idle
    gpu, ora = _run_both((3, 3, 3), 96, 1500, seed=99, launches=3, random_iterations=True, p_expand=20.0,
                         partner_layout=layout)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
    assert len(set(int(x) for x in ora["it"])) > 10  # iteration counts really differ per walker


@pytest.mark.parametrize("walkers", [1, 33, 64, 200])
def test_one_lane_layout_ragged_pools(walkers):
    # pools that end inside a block and inside a wave of the one-lane layout
    gpu, ora = _run_both((2, 3, 3), walkers, 1200, seed=404, launches=2, plus=100, p_expand=80.0, partner_layout=3)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


@pytest.mark.parametrize("n,walkers,iterations", [((2, 2, 2), 128, 2500), ((3, 3, 3), 128, 3000),
                                                 ((4, 4, 4), 64, 1500)])
def test_integer_walk_bit_identical_to_oracle(n, walkers, iterations):
    gpu, ora = _run_both(n, walkers, iterations, seed=4242, ring=O.RING_ZT)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_integer_walk_rare_paths_bit_identical():
    gpu, ora = _run_both((3, 3, 3), 128, 2000, seed=8, launches=2, plus=150, p_expand=200.0, p_reduce=50.0,
                         ring=O.RING_ZT)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
    assert (gpu["cur_s"] != 0).any()  # negative coefficients really occur


def test_walk_from_loaded_states_bit_identical():
    # start from Strassen^2 (rank 49, no flips) split into flippable states
    walkers, iterations, seed = 64, 1500, 2024
    base = O.naive_records(4, 4, 4, walkers)
    p0 = O.make_params(iterations=700, plus_iterations=100, p_expand=300.0)
    start = O.walk_records(base, p0, seed=1, nthreads=8)
    with fg.Engine(walkers, seed, max_iterations=iterations, plus_iterations=100, expand_probability=3.0) as e:
        e.init_naive(4, 4, 4)
        e.write_walkers(start)
        assert first_mismatch(e.read_walkers(), start) is None
        assert e.walk(iterations) == 0
        gpu = e.read_walkers()
    ora = O.walk_records(start, O.make_params(iterations=iterations, plus_iterations=100, p_expand=3.0),
                         seed, nthreads=8)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_gpu_brent_check_matches_oracle():
    rng = np.random.default_rng(0)
    blobs, want = [], []
    for sch in (S.strassen(), S.strassen_squared()):
        b = S.to_blob(*sch)
        blobs.append(b)
        want.append(True)
        for _ in range(20):
            c = b.copy()
            pos = int(rng.integers(4, len(c)))
            c[pos] = 1 - abs(int(c[pos]))
            blobs.append(c)
            want.append(O.validate(c, O.RING_Z2) == 1)
    for n in [(2, 2, 2), (3, 3, 3), (3, 4, 5), (4, 4, 4), (5, 5, 5)]:
        blobs.append(O.naive_blob(*n))
        want.append(True)
    got = fg.verify(blobs)
    assert list(got) == want


def test_gpu_integer_brent_check_matches_oracle():
    rng = np.random.default_rng(1)
    blobs, want = [], []
    for sch in (S.strassen(), S.strassen_squared()):
        b = S.to_blob(*sch)
        blobs.append(b)
        want.append(True)
        for _ in range(20):
            c = b.copy()
            pos = int(rng.integers(4, len(c)))
            c[pos] = int(rng.choice([x for x in (-1, 0, 1) if x != int(c[pos])]))
            blobs.append(c)
            want.append(O.validate(c, O.RING_ZT) == 1)
    got = fg.verify(blobs, ring=fg.RING_ZT)
    assert list(got) == want


def test_best_scheme_is_valid_and_host_reverified():
    with fg.Engine(256, 5, max_iterations=4000, plus_iterations=4000) as e:
        e.init_naive(3, 3, 3)
        assert e.walk() == 0
        blob, rank, walker = e.best()
        assert rank == int(blob[3]) and rank <= 27
        assert O.validate(blob) == 1
        n, U, V, W = S.from_blob(blob)
        assert S.brent_ok(n, U, V, W, "Z2")
