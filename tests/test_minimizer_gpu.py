"""GPU complexity minimizer (fg_init_minimizer / fg_minimize /
fg_broadcast_best) vs the oracle's minimize mode: bit-identical walker
records through two launches with the reference's re-broadcast between them
(src/entities/complexity_minimizer.cu:130-133)."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
from minimizer_util import (broadcast, minimizer_start, record_blob, record_complexity, rebroadcast_targets,
                            walked_schemes)
from walk_compare import first_mismatch

pytestmark = pytest.mark.gpu


def _blob_of(schemes):
    return np.concatenate([np.array([len(schemes)], dtype=np.int32)] + list(schemes))


@pytest.mark.parametrize("ring,n,walkers,iterations,bits", [
    (O.RING_Z2, (3, 3, 3), 1024, 3000, 0),
    (O.RING_Z2, (4, 4, 4), 256, 2000, 0),
    (O.RING_Z2, (3, 3, 3), 256, 2000, 16),
    (O.RING_ZT, (3, 3, 3), 512, 3000, 0),
    (O.RING_ZT, (4, 4, 4), 256, 2000, 0),
])
def test_minimizer_bit_identical_to_oracle(ring, n, walkers, iterations, bits):
    print(f"[minimizer] ring={ring} n={n} walkers={walkers} iterations={iterations} bits={bits}", flush=True)
    schemes = walked_schemes(n, 24, 2000, 5, ring)
    seed = 2024
    with fg.Engine(walkers, seed, ring=ring, max_iterations=iterations, check_reduce=False, word_bits=bits) as e:
        e.init_minimizer(_blob_of(schemes))
        init = e.read_walkers()
        # initializeKernel: round-robin copies, best = start, best complexity = global minimum
        cx0 = min(O.complexity(b, ring) for b in schemes)
        for w in (0, 1, len(schemes), walkers - 1):
            assert np.array_equal(record_blob(init[w], ring), schemes[w % len(schemes)])
        assert (init["best_rank"] == cx0).all()
        start, cx = minimizer_start(init, ring)
        assert cx == cx0 and first_mismatch(init, start) is None

        p = O.make_params(iterations=iterations, mode=1, check_reduce=False)
        assert e.minimize() == 0
        gpu = e.read_walkers()
        ora = O.walk_records(start, p, seed, ring=ring, nthreads=8)
        assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)
        best, cur = e.complexities()
        assert np.array_equal(best, ora["best_rank"])
        assert all(cur[w] == record_complexity(ora[w], ring) for w in range(0, walkers, 7))

        # report(): re-broadcast of the best walker, then a second launch
        indices = sorted(range(walkers), key=lambda w: (int(best[w]), w))
        dst = rebroadcast_targets(indices, 1, 4)
        e.broadcast_best(indices[0], dst)
        assert e.minimize() == 0
        gpu2 = e.read_walkers()
        ora2 = O.walk_records(broadcast(ora, indices[0], dst), p, seed, ring=ring, nthreads=8)
        assert first_mismatch(gpu2, ora2) is None, first_mismatch(gpu2, ora2)
        assert (ora2["best_rank"] < cx0).any()

        blob, bc, bw = e.best()
        assert bc == int(ora2["best_rank"].min()) and O.validate(blob, ring) == 1
        assert O.complexity(blob, ring) == bc


def test_minimizer_rejects_invalid_and_mixed_input():
    print("[minimizer] invalid input", flush=True)
    good = O.naive_blob(2, 2, 2)
    bad = good.copy()
    bad[4] ^= 1
    with fg.Engine(64, 1, check_reduce=False) as e:
        with pytest.raises(fg.FgError):
            e.init_minimizer(_blob_of([good, bad]))
        with pytest.raises(fg.FgError):
            e.init_minimizer(_blob_of([good, O.naive_blob(2, 2, 3)]))
        with pytest.raises(fg.FgError):
            e.minimize()  # no pool
        e.init_naive(2, 2, 2)
        with pytest.raises(fg.FgError):
            e.minimize()  # a walk pool
