"""GPU-recorded flip traces (fg_walk_trace) through the C ABI.  Each walker's
trace -- every flip, plus, split and tryReduce the walk kernel applied, with
the arguments its random draws chose -- is replayed by the oracle from the
walker's scheme before the launch (fgo_blob_replay: explicit operator calls,
no random draws) and must end on the GPU's scheme after it, term order
included.  This checks the GPU's operators on the walk's own trajectories
without relying on the shared Philox streams.  The traced launch must also
take the plain launch's trajectory (the oracle's walk from the same state)."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
import schemes_np as S
from minimizer_util import record_blob, records_from_blob
from walk_compare import first_mismatch

pytestmark = pytest.mark.gpu

# (format, ring, partner_layout, word_bits): Z2 class masks, 16-bit and
# 32-bit words with partner lists, and the integer ring
CASES = [
    ((3, 3, 3), O.RING_Z2, 2, 0), ((3, 3, 3), O.RING_Z2, 1, 16), ((3, 3, 3), O.RING_Z2, 1, 32),
    ((2, 3, 4), O.RING_Z2, 2, 0), ((4, 4, 4), O.RING_Z2, 0, 16), ((4, 4, 4), O.RING_Z2, 0, 32),
    ((3, 3, 3), O.RING_ZT, 0, 0), ((4, 4, 4), O.RING_ZT, 0, 0),
    ("strassen", O.RING_ZT, 0, 0), ("strassen", O.RING_Z2, 2, 0),
    ("strassen_squared", O.RING_ZT, 0, 0), ("strassen_squared", O.RING_Z2, 0, 16),
    ((3, 3, 3), O.RING_Z2, 3, 0), ("strassen", O.RING_Z2, 3, 0),
]


def known_start(name, ring, walkers):
    """Walkers on Strassen's scheme (2x2x2, rank 7) or its Kronecker square
    (4x4x4, rank 49): below the naive rank, so plus and split have room."""
    blob = S.to_blob(*(S.strassen() if name == "strassen" else S.strassen_squared()))
    if ring == O.RING_ZT:
        blob = O.replay(blob, [[5]], ring=ring)  # fixSigns, as the readers do
    return tuple(int(x) for x in blob[:3]), records_from_blob(blob, walkers, ring=ring)


@pytest.mark.parametrize("n,ring,layout,word_bits", CASES)
def test_gpu_trace_replays_to_the_launch_result(n, ring, layout, word_bits):
    walkers, iterations, seed = 128, 600, 77
    print(f"[trace] n={n} ring={ring} layout={layout} word_bits={word_bits}", flush=True)
    kw = dict(ring=ring, max_iterations=iterations, plus_iterations=60, expand_probability=80.0,
              reduce_probability=200.0, zt_visit_cap=16, word_bits=word_bits, partner_layout=layout)
    with fg.Engine(walkers, seed, **kw) as e:
        if isinstance(n, str):
            n, recs = known_start(n, ring, walkers)
            e.init_naive(*n)
            e.write_walkers(recs)
        else:
            e.init_naive(*n)
        assert e.walk(iterations) == 0  # start the trace from walked schemes
        start = e.read_walkers()
        cap = 6 * iterations
        ops, counts, bad = e.walk_trace(iterations, cap)
        assert bad == 0
        out = e.read_walkers()
    p = O.make_params(iterations=iterations, plus_iterations=60, p_expand=80.0, p_reduce=200.0)
    ora = O.walk_records(start, p, seed, ring=ring, nthreads=8)
    assert first_mismatch(out, ora) is None, first_mismatch(out, ora)
    assert (counts <= cap).all() and (counts > 0).all()
    for w in range(walkers):
        got = O.replay(record_blob(start[w], ring), ops[w, :counts[w]], ring=ring)
        assert np.array_equal(got, record_blob(out[w], ring)), f"walker {w}: replay of {counts[w]} operators differs"
    codes = set(np.unique(np.concatenate([ops[w, :counts[w], 0] for w in range(walkers)])).tolist())
    assert codes <= {0, 1, 2, 6}
    assert 0 in codes and 6 in codes, codes
    if (start["m"] < np.prod(start["n"], axis=1)).any():  # plus and split need room below the naive rank
        assert 1 in codes or 2 in codes, codes


def test_gpu_trace_cut_at_capacity():
    """A capacity below the operator count stores the first cap operators and
    still reports the full count; the walk itself is unchanged."""
    with fg.Engine(64, 5, max_iterations=200, plus_iterations=60) as e:
        e.init_naive(3, 3, 3)
        start = e.read_walkers()
        full_ops, full_n, _ = e.walk_trace(200, 1000)
    with fg.Engine(64, 5, max_iterations=200, plus_iterations=60) as e:
        e.init_naive(3, 3, 3)
        ops, n, bad = e.walk_trace(200, 50)
        assert bad == 0
        out = e.read_walkers()
    assert np.array_equal(n, full_n) and (n > 50).all()
    assert np.array_equal(ops, full_ops[:, :50])
    p = O.make_params(iterations=200, plus_iterations=60)
    assert first_mismatch(out, O.walk_records(start, p, 5, nthreads=8)) is None
