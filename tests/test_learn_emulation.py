"""The learn-mode walk kernel's lane code (proposed first flip, logged pair,
lowest rank of the launch), run on the host, matches the oracle's forced
walk bit for bit."""
import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
from learn_util import random_forced, record_pairs
from walk_compare import first_mismatch


@pytest.mark.parametrize("n,iters", [((2, 2, 2), 300), ((3, 3, 3), 500), ((4, 4, 4), 400)])
def test_forced_launch_matches_oracle(n, iters):
    rng = np.random.default_rng(sum(n))
    recs = O.walk_records(O.naive_records(*n, 64), O.make_params(iterations=700, plus_iterations=300), 5, nthreads=8)
    forced = random_forced(recs, rng)
    p = O.make_params(iterations=iters, plus_iterations=300)
    emu, err, lm, mm = E.walk(recs, p, 9, forced=forced, learn=True)
    ora, log, omm = O.walk_records_forced(recs, p, 9, forced)
    assert err == 0
    assert first_mismatch(emu, ora) is None, first_mismatch(emu, ora)
    assert np.array_equal(lm, log) and np.array_equal(mm, omm)
    # a proposed pair that is a flip is the one taken
    for w in range(len(recs)):
        if forced[w, 0] >= 0:
            assert tuple(log[w, :3]) == tuple(forced[w, :3])
        assert log[w, 0] < 0 or tuple(log[w, :3]) in record_pairs(recs[w])
        assert log[w, 3] == recs[w]["m"] and mm[w] <= recs[w]["m"]


def test_no_proposal_is_the_plain_walk():
    recs = O.walk_records(O.naive_records(3, 3, 3, 64), O.make_params(iterations=500), 1, nthreads=8)
    p = O.make_params(iterations=400)
    plain = O.walk_records(recs, p, 4)
    ora, log, mm = O.walk_records_forced(recs, p, 4, None)
    assert first_mismatch(plain, ora) is None
    emu, err, lm, emm = E.walk(recs, p, 4, forced=None, learn=True)
    assert first_mismatch(emu, plain) is None and np.array_equal(lm, log)
