"""The scorer's CPU oracle (oracle/scorer_ref.py): its backward pass is
pinned by central finite differences of its own forward pass, its bfloat16
rounding by exact cases, and its term features by the bit layout."""
import numpy as np

from learn_util import SR, random_batch


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3])
    r = SR.bf16(x)
    assert r[0] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[4] == -2.5
    assert r[1] == 1.0                      # tie to even
    assert r[3] == 1.0 + 2 ** -6            # tie to even upward
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_features_are_the_factor_bits():
    words = np.zeros((1, 3, 4), dtype=np.uint32)
    words[0, 0, 1] = 0b1001
    words[0, 2, 0] = 0b10
    X, mask = SR.features(words, [2], (4, 4, 4))
    assert X.shape == (1, 4, 12) and list(mask[0]) == [1, 1, 0, 0]
    assert list(X[0, 1, :4]) == [1, 0, 0, 1] and X[0, 1, 4:].sum() == 0
    assert X[0, 0, 8 + 1] == 1 and X[0, 0].sum() == 1


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    nn = (4, 4, 4)
    F, M = sum(nn), 6
    layout, n = SR.make_layout(F, M)
    p = SR.init_params(F, 5)
    words, m, pairs, y = random_batch(rng, 5, M, nn)
    loss, g, _ = SR.loss_and_grads(p, layout, words, m, pairs, y, nn, exact=True)
    off = list(layout[4:]) + [n]
    h = 1e-6
    for k in range(12):
        a, b = off[k], off[k + 1]
        for idx in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            q = p.copy()
            q[idx] += h
            lp = SR.loss_and_grads(q, layout, words, m, pairs, y, nn, exact=True)[0]
            q[idx] -= 2 * h
            lm = SR.loss_and_grads(q, layout, words, m, pairs, y, nn, exact=True)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 + 1e-4 * abs(fd), (k, idx, fd, g[idx])
    assert np.abs(g).sum() > 0
