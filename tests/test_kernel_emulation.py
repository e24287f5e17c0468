"""The walk kernel's own lane code, run on the host (tests/emu), must match
the oracle bit-for-bit.  This is the CPU-side gate before any GPU run: it
exercises the partner-class bookkeeping, selection and rare paths of the
exact source the GPU compiles."""
import pytest

import emu_ffi as E
import oracle_ffi as O
from walk_compare import first_mismatch


def _check(n, walkers, iterations, seed, launches=1, plus=8000, p_expand=0.0, random_iterations=False,
           check_reduce=True, ring=O.RING_Z2, p_reduce=0.0, layout="auto"):
    start = O.naive_records(*n, walkers, ring=ring)
    p = O.make_params(iterations=iterations, plus_iterations=plus, p_expand=p_expand, p_reduce=p_reduce,
                      random_iterations=random_iterations, check_reduce=check_reduce)
    emu, err = E.walk(start, p, seed, launches=launches, ring=ring, layout=layout)
    ora = O.walk_records(start, p, seed, launches=launches, nthreads=8, ring=ring)
    assert err == 0
    mism = first_mismatch(emu, ora)
    assert mism is None, mism
    return ora


@pytest.mark.parametrize("n,walkers,iterations", [
    ((2, 2, 2), 64, 2000), ((2, 2, 3), 64, 2000), ((3, 3, 3), 64, 3000), ((2, 3, 4), 64, 2000),
    ((4, 4, 4), 64, 1500), ((3, 4, 5), 32, 1000),
])
def test_emulated_kernel_matches_oracle(n, walkers, iterations):
    ora = _check(n, walkers, iterations, seed=12345)
    assert (ora["total_flips"] > 0).all()


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (3, 3, 4)])
def test_emulated_kernel_plus_expand_paths(n):
    _check(n, 64, 2000, seed=7, launches=2, plus=150, p_expand=200.0)


# schemes of up to 32 terms run the class-mask lanes by default; these run
# the same cases on the partner-list lanes (16- and 32-bit factor words) too
@pytest.mark.parametrize("layout", ["masks", "half", "wide", "solo"])
@pytest.mark.parametrize("n", [(2, 2, 2), (2, 3, 4), (3, 3, 3)])
def test_emulated_layouts_match_oracle(n, layout):
    _check(n, 64, 2000, seed=17, launches=2, plus=150, p_expand=200.0, layout=layout)


def test_emulated_kernel_random_iterations():
    _check((3, 3, 3), 64, 1200, seed=99, launches=3, random_iterations=True, p_expand=20.0)


@pytest.mark.parametrize("layout", ["masks", "half", "solo"])
def test_emulated_kernel_without_reductions(layout):
    # complexity-minimizer mode: tryFlip(checkReduce = false)
    ora = _check((3, 3, 3), 64, 2000, seed=3, check_reduce=False, layout=layout)
    assert (ora["m"] >= 27).all() or (ora["m"] < 27).any()


@pytest.mark.parametrize("layout", ["masks", "half", "solo"])
def test_emulated_kernel_reduce_probability(layout):
    _check((3, 3, 3), 64, 1500, seed=11, p_reduce=300.0, plus=200, p_expand=50.0, layout=layout)


@pytest.mark.parametrize("n,walkers,iterations", [
    ((2, 2, 2), 64, 2000), ((2, 2, 3), 64, 2000), ((3, 3, 3), 64, 2500), ((4, 4, 4), 64, 1500),
])
def test_emulated_integer_kernel_matches_oracle(n, walkers, iterations):
    ora = _check(n, walkers, iterations, seed=4242, ring=O.RING_ZT)
    assert (ora["total_flips"] > 0).all()


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3)])
def test_emulated_integer_kernel_plus_expand_paths(n):
    _check(n, 64, 2000, seed=8, launches=2, plus=150, p_expand=200.0, ring=O.RING_ZT, p_reduce=50.0)


@pytest.mark.parametrize("n", [(2, 2, 2), (3, 3, 3), (4, 4, 4)])
def test_emulated_wide_words_match_oracle(n):
    # the 32-bit-word lane code, which formats with wider factors use,
    # on formats that otherwise run with 16-bit words
    start = O.naive_records(*n, 64)
    p = O.make_params(iterations=1500, plus_iterations=200, p_expand=100.0)
    emu, err = E.walk(start, p, 77, launches=2, wide=True)
    ora = O.walk_records(start, p, 77, launches=2, nthreads=8)
    assert err == 0
    assert first_mismatch(emu, ora) is None, first_mismatch(emu, ora)


@pytest.mark.parametrize("layout", ["masks", "half", "wide", "solo"])
@pytest.mark.parametrize("n", [(3, 3, 3), (4, 4, 4)])
def test_emulated_lookahead_matches_oracle(n, layout):
    # reduction-seeking lookahead proposals every 2nd iteration, 3 candidates
    if n == (4, 4, 4) and layout in ("masks", "solo"):
        pytest.skip("class masks hold at most 32 terms")
    start = O.naive_records(*n, 32)
    p = O.make_params(iterations=1500, plus_iterations=300, p_expand=50.0, look_period=2, look_candidates=3)
    emu, err = E.walk(start, p, 21, launches=2, layout=layout)
    ora = O.walk_records(start, p, 21, launches=2, nthreads=8)
    assert err == 0
    assert first_mismatch(emu, ora) is None, first_mismatch(emu, ora)
    if n == (3, 3, 3):  # the proposals do find reductions
        base = O.walk_records(start, O.make_params(iterations=1500, plus_iterations=300, p_expand=50.0), 21,
                              launches=2, nthreads=8)
        assert int(ora["m"].sum()) < int(base["m"].sum())
