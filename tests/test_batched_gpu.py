"""Several formats in one persistent launch (config C3 shape, scaled down):
walker allocation, bit-identical trajectories against the oracle, HIP-graph
replay and the grid autotuner."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
from walk_compare import first_mismatch

pytestmark = pytest.mark.gpu

C3_FORMATS = [(2, 2, 3), (2, 3, 4), (3, 3, 4), (3, 4, 5), (4, 4, 4)]
C3_WEIGHTS = [1, 2, 4, 8, 16]


def expected_counts(W, weights):
    total = sum(weights)
    exact = [W * w / total for w in weights]
    cnt = [int(x) for x in exact]
    order = sorted(range(len(weights)), key=lambda f: -(exact[f] - cnt[f]))
    q = 0
    while sum(cnt) < W:
        cnt[order[q % len(order)]] += 1
        q += 1
    return cnt


def test_batched_allocation_and_parity():
    W, iters, seed = 1000, 1500, 77
    print(f"[batched] W={W} iters={iters}", flush=True)
    with fg.Engine(W, seed, max_iterations=iters, plus_iterations=300, expand_probability=10.0) as e:
        e.init_batched(C3_FORMATS, C3_WEIGHTS)
        start = e.read_walkers()
        counts = expected_counts(W, C3_WEIGHTS)
        got = []
        for f in C3_FORMATS:
            got.append(int(np.all(start["n"] == np.array(f), axis=1).sum()))
        assert got == counts
        assert (start["m"] == start["n"].prod(axis=1)).all()
        assert e.walk(iters) == 0
        assert e.walk(iters) == 0
        gpu = e.read_walkers()
    p = O.make_params(iterations=iters, plus_iterations=300, p_expand=10.0)
    ora = O.walk_records(start, p, seed, launches=2, nthreads=8)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_batched_one_lane_layout():
    # formats of up to 32 terms in the one-lane class-mask layout: a wave
    # serves two queue blocks, which may hold different formats
    W, iters, seed = 500, 1200, 78
    fmts, weights = [(2, 2, 3), (2, 3, 4), (3, 3, 3)], [1, 2, 4]
    print(f"[batched-solo] W={W} iters={iters}", flush=True)
    with fg.Engine(W, seed, max_iterations=iters, plus_iterations=300, expand_probability=10.0,
                   partner_layout=3) as e:
        e.init_batched(fmts, weights)
        start = e.read_walkers()
        assert e.walk(iters) == 0
        assert e.walk(iters) == 0
        gpu = e.read_walkers()
    p = O.make_params(iterations=iters, plus_iterations=300, p_expand=10.0)
    ora = O.walk_records(start, p, seed, launches=2, nthreads=8)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_graph_replay_matches_plain_launches():
    W, iters, seed = 512, 800, 5
    out = []
    for use_graph in (False, True):
        print(f"[graph] use_graph={use_graph}", flush=True)
        with fg.Engine(W, seed, max_iterations=iters, plus_iterations=200, use_graph=use_graph) as e:
            e.init_batched(C3_FORMATS[:3], [1, 1, 2])
            for _ in range(3):
                assert e.walk(iters) == 0
            out.append(e.read_walkers())
    assert first_mismatch(out[1], out[0]) is None


@pytest.mark.parametrize("n,W,use_graph", [((3, 3, 3), 2048, False), ((3, 3, 3), 2048, True),
                                            ((4, 4, 4), 512, True)])
def test_autotune_keeps_results_exact(n, W, use_graph):
    """The autotuner times every layout the pool fits (class masks on lane
    pairs or one lane per walker, 16- and 32-bit partner lists) and grid with ordinary 100-iteration launches; the
    walk then continues in the chosen layout, bit-identical to the oracle
    replaying all those launches."""
    seed = 9
    print(f"[autotune] n={n} W={W} use_graph={use_graph}", flush=True)
    with fg.Engine(W, seed, max_iterations=400, plus_iterations=200, use_graph=use_graph) as e:
        e.init_naive(*n)
        start = e.read_walkers()
        assert e.walk(400) == 0  # a launch before tuning (a cached graph of the first layout)
        g = e.autotune(100)
        layout, trials = e.layout_info()
        assert g >= 1 and len(trials) >= 1
        assert layout in set(int(x) for x in trials[:, 0])
        expect_layouts = {0, 2} | ({3, 4} if n[0] * n[1] * n[2] <= 32 else set())
        assert set(int(x) for x in trials[:, 0]) == expect_layouts
        # each trial launched its own grid: every (layout, grid) pair is distinct
        assert len({(int(a), int(b)) for a, b, _ in trials}) == len(trials)
        assert e.walk(400) == 0
        gpu = e.read_walkers()
        c = e.counters()
    assert (gpu["it"] == gpu["it"][0]).all()
    iters_done = int(gpu["it"][0])
    assert c["iterations"] == W * iters_done
    assert iters_done == 800 + 100 * len(trials)
    ora = O.walk_records(start, O.make_params(iterations=400, plus_iterations=200), seed, nthreads=8)
    ora = O.walk_records(ora, O.make_params(iterations=100, max_iterations=400, plus_iterations=200), seed,
                         launches=len(trials), nthreads=8)
    ora = O.walk_records(ora, O.make_params(iterations=400, plus_iterations=200), seed, nthreads=8)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
