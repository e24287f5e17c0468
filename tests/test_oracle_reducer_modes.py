"""The additions reducer's random selection modes (pairs_counter.cuh:133-191)
pinned step by step against fixed Philox words, as the greedy-alternative
draw is in test_oracle_reducer.py.  Each case is a small expression set whose
counted pairs are known by hand, so the pair each mode must take follows
from the stream words of (reducer r, component c) -- the scale draw
(additions_reducer.cuh:91, word 0) first, then the mode's own draws -- and
from the reference's rule.  Every case also checks that the words
discriminate: the same draws under a plausible misreading of the rule (a
count instead of count - 1, a factor 1.0 instead of 0.69, count - 1 weights)
pick a different pair for some seeds."""
import numpy as np

import oracle_ffi as O
import reducer_ffi as R

f32 = np.float32
R_, C_ = 3, 1  # reducer index and component of the stream


def words(seed, n=48):
    out, b = [], 0
    while len(out) < n:
        out += [int(x) for x in O.philox([b, 0, R_, 2 + C_], [seed, 0])]
        b += 1
    return out


def uniform(w):
    """curand_uniform of this build: ((x >> 8) + 1) * 2^-24, in (0, 1]."""
    return f32(((w >> 8) + 1) * 2.0 ** -24)


def rows_of(pairs_in_order, nvars):
    """Rows x_i + x_j (one per occurrence), so the counter sees exactly these
    pairs, in this first-seen order."""
    rows = []
    for (i, j), count in pairs_in_order:
        for _ in range(count):
            r = [0] * nvars
            r[i - 1] = r[j - 1] = 1
            rows.append(r)
    return rows


def test_weighted_random_weights_by_count():
    # pairs (1,2) x3 and (3,4) x2: getWeightedRandom (:173-191) takes (1,2)
    # when u * (3 + 2) <= 3; weights count - 1 would need u * 3 <= 2
    rows = rows_of([((1, 2), 3), ((3, 4), 2)], 4)
    differ = 0
    for seed in range(1, 400):
        w = words(seed)
        u = uniform(w[1])
        expect = (1, 2) if f32(u * f32(5)) <= f32(3) else (3, 4)
        res = R.single(rows, mode=4, seed=seed, r=R_, c=C_)
        assert res["vars"][0] == expect, seed
        differ += expect != ((1, 2) if f32(u * f32(3)) <= f32(2) else (3, 4))
    assert differ > 0


def test_greedy_random_draws_among_all_below_the_scale():
    # pairs (1,2) x3 (the top), (3,4) x2, (5,6) x2: getGreedyRandom (:139-143)
    # draws from all three when u < scale, else from the top one
    rows = rows_of([((1, 2), 3), ((3, 4), 2), ((5, 6), 2)], 6)
    order = [(1, 2), (3, 4), (5, 6)]
    branches = set()
    for seed in range(1, 200):
        w = words(seed)
        scale = f32(uniform(w[0]) / f32(2))
        below = uniform(w[1]) < scale
        expect = order[w[2] % 3] if below else order[0]
        res = R.single(rows, mode=2, seed=seed, r=R_, c=C_)
        assert res["vars"][0] == expect, seed
        assert abs(res["scale"] - float(scale)) == 0.0
        branches.add((below, expect))
    assert {b for b, _ in branches} == {True, False} and len(branches) >= 3


def _intersects(a, b):
    return a[0] == b[0] or a[0] == abs(b[1]) or abs(a[1]) == b[0] or abs(a[1]) == abs(b[1])


def _intersections_pick(pairs, draws, scale, factor=0.69, nonint=lambda c: c - 1):
    """getGreedyIntersections (:145-171) over counted pairs in counter order."""
    imax, best = 0, f32(0)
    for i, (pi, ci) in enumerate(pairs):
        s = f32(0)
        for j, (pj, cj) in enumerate(pairs):
            if i == j:
                continue
            if not _intersects(pi, pj):
                s = f32(s + f32(nonint(cj)))
            else:
                x = next(draws)
                s = f32(float(s) + (0.0 if x % 2 else factor * float(cj)))
        score = f32(f32(ci - 1) + f32(scale * s))
        if score > best:
            best, imax = score, i
    return pairs[imax][0]


def test_greedy_intersections_scores():
    # four pairs seen twice each; (1,6), (2,6) and (6,7) intersect pairwise,
    # (4,5) meets none.  A non-intersecting pair adds count - 1, an
    # intersecting one adds 0.69 * count when its coin is even.
    pairs = [((1, 6), 2), ((2, 6), 2), ((4, 5), 2), ((6, 7), 2)]
    rows = rows_of(pairs, 7)
    wrong_factor = wrong_count = 0
    for seed in range(1, 120):
        w = words(seed)
        scale = f32(uniform(w[0]) / f32(2))
        expect = _intersections_pick(pairs, iter(w[1:]), scale)
        res = R.single(rows, mode=3, seed=seed, r=R_, c=C_)
        assert res["vars"][0] == expect, seed
        wrong_factor += expect != _intersections_pick(pairs, iter(w[1:]), scale, factor=1.0)
        wrong_count += expect != _intersections_pick(pairs, iter(w[1:]), scale, nonint=lambda c: c)
    assert wrong_factor >= 10 and wrong_count >= 10
