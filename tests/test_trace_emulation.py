"""Flip traces recorded by the walk kernels' own lane code, run on the host
(tests/emu).  A traced launch writes, per walker, every operator it applies
(flip, plus, split, tryReduce) with the arguments its random draws chose.
The oracle replays each walker's trace from the walker's scheme before the
launch (fgo_blob_replay: explicit operator calls, no random draws); the
replay must end on the launch's scheme, term order included.  The traced
launch must also take the trajectory of the oracle's walk."""
import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
import schemes_np as S
from minimizer_util import record_blob, records_from_blob
from walk_compare import first_mismatch

CASES = [
    ((2, 2, 2), O.RING_Z2, "wide"), ((2, 2, 2), O.RING_Z2, "half"), ((2, 2, 2), O.RING_Z2, "masks"),
    ((3, 3, 3), O.RING_Z2, "wide"), ((3, 3, 3), O.RING_Z2, "half"), ((3, 3, 3), O.RING_Z2, "masks"),
    ((4, 4, 4), O.RING_Z2, "half"),
    ((2, 2, 2), O.RING_ZT, "auto"), ((3, 3, 3), O.RING_ZT, "auto"),
    ("strassen", O.RING_ZT, "auto"), ("strassen", O.RING_Z2, "masks"),
    ("strassen_squared", O.RING_ZT, "auto"), ("strassen_squared", O.RING_Z2, "half"),
    ((3, 3, 3), O.RING_Z2, "solo"), ("strassen", O.RING_Z2, "solo"),
]


def known_start(name, ring, walkers):
    """Walkers on Strassen's scheme (2x2x2, rank 7) or its Kronecker square
    (4x4x4, rank 49): below the naive rank, so plus and split have room."""
    blob = S.to_blob(*(S.strassen() if name == "strassen" else S.strassen_squared()))
    if ring == O.RING_ZT:
        blob = O.replay(blob, [[5]], ring=ring)  # fixSigns, as the readers do
    return records_from_blob(blob, walkers, ring=ring)


def traced_launch(n, ring, layout, walkers=64, iterations=300, seed=23):
    # a small plateau, frequent expansions and reductions: every operator shows up
    p = O.make_params(iterations=iterations, plus_iterations=40, p_expand=60.0, p_reduce=150.0)
    if isinstance(n, str):
        start = known_start(n, ring, walkers)
    else:
        pre = O.make_params(iterations=5000, plus_iterations=40, p_expand=60.0, p_reduce=150.0)
        start = O.walk_records(O.naive_records(*n, walkers, ring=ring), pre, seed, ring=ring, nthreads=8)
    cap = 6 * iterations
    out, err, ops, counts = E.walk(start, p, seed, ring=ring, layout=layout, trace_cap=cap)
    assert err == 0
    ora = O.walk_records(start, p, seed, ring=ring, nthreads=8)
    return start, out, ora, ops, counts, cap


@pytest.mark.parametrize("n,ring,layout", CASES)
def test_trace_replays_to_the_launch_result(n, ring, layout):
    start, out, ora, ops, counts, cap = traced_launch(n, ring, layout)
    assert first_mismatch(out, ora) is None, first_mismatch(out, ora)
    assert (counts <= cap).all() and (counts > 0).all()
    for w in range(len(start)):
        before, after = record_blob(start[w], ring), record_blob(out[w], ring)
        got = O.replay(before, ops[w, :counts[w]], ring=ring)
        assert np.array_equal(got, after), f"walker {w}: replay of {counts[w]} operators differs"
    codes = set(np.unique(np.concatenate([ops[w, :counts[w], 0] for w in range(len(start))])).tolist())
    assert codes <= {0, 1, 2, 6}
    assert 0 in codes and 6 in codes, codes
    naive = np.prod(start["n"], axis=1)
    if (start["m"] < naive).any():  # plus and split need room below the naive rank
        assert 1 in codes or 2 in codes, codes


def test_a_dropped_operator_breaks_the_replay():
    """The replay check is not vacuous: leaving out the last recorded flip,
    or swapping its two indices, lands on another scheme."""
    start, out, _, ops, counts, _ = traced_launch((3, 3, 3), O.RING_Z2, "masks", walkers=8, iterations=120)
    w = int(np.argmax(counts))
    seq = ops[w, :counts[w]]
    flips = np.flatnonzero(seq[:, 0] == 0)
    after = record_blob(out[w], O.RING_Z2)
    dropped = np.delete(seq, flips[-1], axis=0)
    assert not np.array_equal(O.replay(record_blob(start[w]), dropped), after)
    swapped = seq.copy()
    q = flips[-1]
    swapped[q, 4], swapped[q, 5] = seq[q, 5], seq[q, 4]
    assert not np.array_equal(O.replay(record_blob(start[w]), swapped), after)
