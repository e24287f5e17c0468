"""The oracle's additions reducer (a restatement of
src/entities/additions_reducer.cuh and pairs_counter.cuh) pinned by hand-
worked greedy reductions, the random-draw order of its greedy-alternative
mode, and, for every mode, the invariant that substituting the fresh
variables back gives the input expressions."""
import numpy as np
import pytest

import oracle_ffi as O
import reducer_ffi as R
import schemes_np as S


def test_greedy_shares_the_first_most_frequent_pair():
    # x1+x2+x3, x1+x2, x2+x3: pairs in first-seen order (1,2):2 (1,3):1 (2,3):2;
    # greedy takes (1,2) as x4: x3+x4, x4, x2+x3; then no pair repeats
    res = R.single([[1, 1, 1], [1, 1, 0], [0, 1, 1]], mode=0)
    assert res["naive"] == 4
    assert res["vars"] == [(1, 2)]
    assert res["expr"] == [[3, 4], [4], [2, 3]]
    assert res["additions"] == 3 and res["fresh"] == 1


def test_signs_are_shared_with_negation():
    # x1-x2 and -x1+x2 share (1,-2): one fresh variable and its negation
    res = R.single([[1, -1], [-1, 1]], mode=0)
    assert res["vars"] == [(1, -2)]
    assert res["expr"] == [[3], [-3]]
    assert res["additions"] == 1
    # x1+x2 and x1-x2 share nothing
    res = R.single([[1, 1], [1, -1]], mode=0)
    assert res["fresh"] == 0 and res["additions"] == 2


def test_greedy_chains_fresh_variables():
    # four copies of x1+x2+x3+x4: (1,2) first -> x5, then (3,5) ... each step
    # takes the first most frequent pair; three fresh variables remain
    res = R.single([[1, 1, 1, 1]] * 4, mode=0)
    assert res["vars"] == [(1, 2), (3, 4), (5, 6)]
    assert res["expr"] == [[7]] * 4
    assert res["additions"] == 3 and res["naive"] == 12


def test_greedy_alternative_draws_after_the_scale():
    # two most frequent pairs, (1,2) and (3,4), both seen twice; word 0 of
    # the stream is the scale draw, word 1 picks among the top pairs
    rows = [[1, 1, 0, 0], [1, 1, 0, 0], [0, 0, 1, 1], [0, 0, 1, 1]]
    for seed in range(1, 9):
        w = O.philox([0, 0, 7, 2 + 1], [seed, 0])
        res = R.single(rows, mode=1, seed=seed, r=7, c=1)
        assert res["scale"] == np.float32(((w[0] >> 8) + 1) * 2.0 ** -24) / np.float32(2)
        first = [(1, 2), (3, 4)][w[1] % 2]
        assert res["vars"][0] == first
        assert sorted(res["vars"]) == [(1, 2), (3, 4)]


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6])
def test_every_mode_preserves_the_expressions(mode):
    rng = np.random.default_rng(100 + mode)
    for trial in range(12):
        ne, nv = int(rng.integers(3, 30)), int(rng.integers(3, 16))
        rows = rng.choice([-1, 0, 0, 1], size=(ne, nv))
        rows[rows.sum(axis=1) == 0, 0] = 1
        res = R.single(rows, mode=mode, seed=trial + 1, r=trial, c=trial % 3)
        assert R.expand(res, nv) == R.rows_as_forms(rows)
        assert res["additions"] == res["fresh"] + sum(len(e) - 1 for e in res["expr"])
        assert res["additions"] <= res["naive"]
        # the loop only stops when no pair occurs twice
        seen = {}
        for e in res["expr"]:
            for a in range(len(e)):
                for b in range(a + 1, len(e)):
                    i, j = e[a], e[b]
                    if abs(i) > abs(j):
                        i, j = j, i
                    if i < 0:
                        i, j = -i, -j
                    seen[(i, j)] = seen.get((i, j), 0) + 1
        assert not seen or max(seen.values()) < 2


def test_strassen_has_no_shared_pair():
    # no signed pair of Strassen's U rows, V rows or W columns occurs twice,
    # so pairwise sharing keeps its 18 additions (a sign slip would share
    # e.g. M1-M5 of c11 with nothing but still count it twice)
    n, U, V, W = S.strassen()
    blob = S.to_blob(n, U, V, W)
    tot_naive = tot = 0
    for c in range(3):
        rows = R.component_rows(blob, c)
        res = R.single(rows, mode=0, c=c, max_fresh=R.MAX_FRESH[c])
        assert R.expand(res, rows.shape[1]) == R.rows_as_forms(rows)
        tot_naive += res["naive"]
        tot += res["additions"]
    assert tot_naive == 18 == O.complexity(blob, O.RING_ZT)
    assert tot == 18


@pytest.mark.parametrize("independent", [True, False])
def test_iteration_streams_and_best_replay(independent):
    n, U, V, W = S.strassen_squared()
    blob = O.replay(S.to_blob(n, U, V, W), [[5]], ring=O.RING_ZT)  # fixSigns as the reader does
    res, words = R.iteration([blob], 24, seed=9, independent=independent)
    assert all(res[0][c]["mode"] == 0 for c in range(3))
    for r in range(24):
        for c in range(3):
            rows = R.component_rows(blob, c)
            assert R.expand(res[r][c], rows.shape[1]) == R.rows_as_forms(rows)
    best = []
    for c in range(3):
        r = min(range(24), key=lambda q: (res[q][c]["additions"], res[q][c]["fresh"]))
        best.append((res[r][c]["fresh"], res[r][c]["vars"]))
    res2, words2 = R.iteration([blob], 24, seed=9, independent=independent, best=best, words=words)
    assert (words2 >= words).all() and (words2[:72:4] > words[:72:4]).all()
    for r in range(24):
        for c in range(3):
            rows = R.component_rows(blob, c)
            assert R.expand(res2[r][c], rows.shape[1]) == R.rows_as_forms(rows)
    # a replayed prefix of the best's fresh variables shows up verbatim
    replayed = [(r, c) for r in range(24) for c in range(3)
                if res2[r][c]["vars"][:1] == best[c][1][:1] and best[c][0] > 0]
    assert replayed
