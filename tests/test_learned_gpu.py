"""Learned proposals on the GPU: the BF16/MFMA scorer's forward and backward
against the float64 oracle (same bfloat16 rounding points), the forced-flip
walk and the trajectory log against the oracle's walk, training on the log,
and the proposals the trained scorer makes."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
from learn_util import SR, random_batch, random_forced, record_pairs
from walk_compare import first_mismatch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.mark.parametrize("count,M,n", [(70, 64, (4, 4, 4)), (37, 27, (3, 3, 3)), (24, 125, (5, 5, 5))])
def test_scorer_forward_backward_match_oracle(count, M, n):
    print(f"[scorer] eval count={count} M={M} n={n}", flush=True)
    nn = (n[0] * n[1], n[1] * n[2], n[2] * n[0])
    rng = np.random.default_rng(count)
    with fg.Scorer(*n, M, seed=3) as sc:
        layout = sc.layout()
        ref_layout, npar = SR.make_layout(sum(nn), M)
        assert list(layout) == list(ref_layout) and sc.nparams == npar
        p = SR.init_params(sum(nn), 11).astype(np.float32)
        sc.set_params(p)
        assert np.array_equal(sc.get_params(), p)
        words, m, pairs, y = random_batch(rng, count, M, nn)
        s_gpu, loss_gpu, g_gpu = sc.eval(words, m, pairs, y)
        loss_ref, g_ref, s_ref = SR.loss_and_grads(p.astype(np.float64), layout, words, m, pairs, y, nn)
        assert _rel(s_gpu, s_ref) < 1e-2
        assert abs(loss_gpu - loss_ref) <= 1e-3 * max(1.0, abs(loss_ref))
        off = list(layout[4:]) + [npar]
        for k in range(12):
            a, b = off[k], off[k + 1]
            assert _rel(g_gpu[a:b], g_ref[a:b]) < 3e-2, (k, _rel(g_gpu[a:b], g_ref[a:b]))
        s2, _, _ = sc.eval(words, m, pairs)
        assert np.array_equal(s2, s_gpu)


@pytest.mark.parametrize("bits", [32, 16])
def test_forced_walk_matches_oracle(bits):
    print(f"[learned] forced walk 4x4x4 bits={bits}", flush=True)
    rng = np.random.default_rng(7)
    W = 256
    with fg.Engine(W, 21, max_iterations=300, plus_iterations=300, word_bits=bits) as e:
        e.init_naive(4, 4, 4)
        assert e.walk(300) == 0
        recs = e.read_walkers()
        forced = random_forced(recs, rng)
        e.set_forced(forced)
        assert np.array_equal(e.get_forced(), forced)
        assert e.walk_forced(300) == 0
        gpu = e.read_walkers()
    ora, log, mm = O.walk_records_forced(recs, O.make_params(iterations=300, plus_iterations=300), 21, forced)
    assert first_mismatch(gpu, ora) is None, first_mismatch(gpu, ora)


def test_trajectory_log_matches_oracle():
    print("[learned] trajectory log", flush=True)
    W, L, H, n = 128, 150, 3, (3, 3, 3)
    p = O.make_params(iterations=L, plus_iterations=400)
    with fg.Engine(W, 5, max_iterations=L, plus_iterations=400) as e:
        e.init_naive(*n)
        e.log_enable(max_batches=8, every=1, horizon=H)
        for _ in range(6):
            assert e.walk(L) == 0
        info = e.log_info()
        assert info["batches"] == 6 - H + 1
        words, m, pairs, y = e.log_read(0, info["rows"], 27)
    # the same launches in the oracle
    recs = O.naive_records(*n, W)
    states, logs, mins = [], [], []
    for _ in range(6):
        states.append(recs.copy())
        recs, log, mm = O.walk_records_forced(recs, p, 5, None)
        logs.append(log)
        mins.append(mm)
    rows = []
    for s in range(6 - H + 1):
        lo = np.min(np.stack(mins[s:s + H]), axis=0)
        for w in range(W):
            if logs[s][w, 0] >= 0:
                rows.append((s, w, tuple(logs[s][w, :3]), int(logs[s][w, 3]), float(lo[w] < logs[s][w, 3])))
    assert len(rows) == info["rows"] and info["positives"] == sum(r[4] for r in rows)
    for q, (s, w, pr, rank, lab) in enumerate(rows):
        assert tuple(pairs[q]) == pr and m[q] == rank and y[q] == lab
        st = states[s][w]
        assert np.array_equal(words[q][:, :rank], st["cur_v"][:, :rank].astype(np.uint32))
        assert pr in record_pairs(st)


def test_window_labels_match_oracle():
    # labels from the first WIN iterations of the logging launch itself
    print("[learned] window labels", flush=True)
    W, L, WIN, n = 128, 150, 20, (3, 3, 3)
    p = O.make_params(iterations=L, plus_iterations=400, min_window=WIN)
    with fg.Engine(W, 6, max_iterations=L, plus_iterations=400) as e:
        e.init_naive(*n)
        e.log_enable(max_batches=8, every=1, horizon=1, window=WIN)
        for _ in range(6):
            assert e.walk(L) == 0
        info = e.log_info()
        assert info["batches"] == 6
        words, m, pairs, y = e.log_read(0, info["rows"], 27)
    recs = O.naive_records(*n, W)
    rows = []
    for s in range(6):
        st = recs.copy()
        recs, log, mm = O.walk_records_forced(recs, p, 6, None)
        for w in range(W):
            if log[w, 0] >= 0:
                rows.append((st[w], tuple(log[w, :3]), int(log[w, 3]), float(mm[w] < log[w, 3])))
    assert len(rows) == info["rows"] and info["positives"] == sum(r[3] for r in rows) > 0
    for q, (st, pr, rank, lab) in enumerate(rows):
        assert tuple(pairs[q]) == pr and m[q] == rank and y[q] == lab
        assert np.array_equal(words[q][:, :rank], st["cur_v"][:, :rank].astype(np.uint32))


def test_training_is_deterministic():
    # fixed-order reductions everywhere: the same log, seeds and steps give
    # the same weights, bit for bit
    print("[learned] deterministic training", flush=True)
    with fg.Engine(2048, 4, max_iterations=100, plus_iterations=2000) as e:
        e.init_naive(4, 4, 4)
        e.log_enable(max_batches=6, every=1, horizon=2)
        for _ in range(7):
            assert e.walk(100) == 0
        out = []
        for _ in range(2):
            with fg.Scorer(4, 4, 4, 64, seed=2) as sc:
                losses = sc.train(e, steps=30, batch=256, lr=2e-3, seed=5)
                out.append((sc.get_params(), losses))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def _auc(scores, labels):
    from scipy.stats import rankdata
    pos = labels > 0.5
    P, N = int(pos.sum()), int((~pos).sum())
    r = rankdata(scores)
    return float((r[pos].sum() - P * (P + 1) / 2) / (P * N))


def _descent_log(seed, W, warm, launches):
    e = fg.Engine(W, seed, max_iterations=1000, plus_iterations=20000)
    e.init_naive(4, 4, 4)
    for _ in range(warm):
        assert e.walk(1000) == 0
    e.log_enable(max_batches=launches, every=1, horizon=1)
    for _ in range(launches):
        assert e.walk(1000) == 0
    return e


def test_trained_network_beats_untrained_on_heldout():
    """What training gains: on flips logged by a fresh uniform walk (another
    seed) at 4x4x4 after 30 launches of descent, labelled 1 when the rank
    falls within the flip's launch, the trained network ranks the positive
    flips better than its untrained initial state (same seed) and than
    uniform scores (AUC 0.5).  Training is deterministic, so the margins are
    reproducible."""
    print("[learned] held-out gain", flush=True)
    e = _descent_log(1000, 16384, 30, 8)
    info = e.log_info()
    assert 0 < info["positives"] < info["rows"]
    with fg.Scorer(4, 4, 4, 64, seed=7) as sc, fg.Scorer(4, 4, 4, 64, seed=7) as sc0:
        sc.train(e, steps=1000, batch=1024, lr=1e-3, seed=11)
        e.close()
        h = _descent_log(2000, 4096, 30, 4)
        n = h.log_info()["rows"]
        words, m, pairs, y = h.log_read(0, n, 64)
        h.close()
        auc_t = _auc(sc.eval(words, m, pairs)[0], y)
        auc_u = _auc(sc0.eval(words, m, pairs)[0], y)
    print(f"[learned] held-out AUC trained {auc_t:.4f} untrained {auc_u:.4f} ({int((y > 0.5).sum())} of {n})", flush=True)
    assert auc_t >= 0.54 and auc_t >= auc_u + 0.03


@pytest.mark.parametrize("n,M,W", [((4, 4, 4), 64, 4096), ((5, 5, 5), 125, 2048)])
def test_training_and_proposals(n, M, W):
    # 5x5x5: 25-coefficient factors in 32-bit words, 125 terms, the chunked
    # proposal path
    print(f"[learned] train + propose n={n}", flush=True)
    with fg.Engine(W, 3, max_iterations=200, plus_iterations=2000) as e, fg.Scorer(*n, M, seed=2) as sc:
        e.init_naive(*n)
        e.log_enable(max_batches=16, every=1, horizon=4)
        for _ in range(12):
            assert e.walk(200) == 0
        info = e.log_info()
        assert info["rows"] >= 8 * W and 0 < info["positives"] < info["rows"]
        losses = sc.train(e, steps=150, batch=512, lr=2e-3, seed=5)
        assert np.isfinite(losses).all()
        assert losses[-20:].mean() < losses[:10].mean()
        sc.propose(e, candidates=8, every=2, round_=1)
        forced = e.get_forced()
        recs = e.read_walkers()
        proposed = 0
        for w in range(0, W, 97):
            if w % 2:
                assert forced[w, 0] == -1
            elif forced[w, 0] >= 0:
                proposed += 1
                assert tuple(forced[w, :3]) in record_pairs(recs[w])
        assert proposed > 0
        bad = e.walk_learned(sc, 200, candidates=8, every=2, round_=2)
        assert bad == 0


@pytest.mark.parametrize("n,M", [((4, 4, 4), 64), ((3, 3, 3), 27), ((5, 5, 5), 125)])
def test_fused_proposals_match_oracle_scores(n, M):
    # the fused proposal kernel (whole network in one workgroup per walker;
    # schemes of up to 64 terms) or the chunked path (5x5x5: 125 terms, 75
    # factor bits) scores every candidate as the float64 oracle does at the
    # same bf16 rounding points, and proposes the first best-scoring candidate
    print(f"[learned] proposals n={n}", flush=True)
    W, K, every = 512, 8, 2
    nn = (n[0] * n[1], n[1] * n[2], n[2] * n[0])
    with fg.Engine(W, 9, max_iterations=300, plus_iterations=300, max_rank=M, word_bits=32) as e, \
            fg.Scorer(*n, M, seed=5) as sc:
        e.init_naive(*n)
        for _ in range(2):
            assert e.walk(300) == 0
        p = SR.init_params(sum(nn), 13).astype(np.float32)
        sc.set_params(p)
        sc.propose(e, candidates=K, every=every, round_=3)
        walkers, pairs, scores = sc.last_proposals(W, K)
        forced = e.get_forced()
        recs = e.read_walkers()
        layout = sc.layout()
    assert len(walkers) == (W + every - 1) // every
    assert np.array_equal(walkers, np.arange(0, W, every))
    rows, ref_idx = [], []
    words = np.zeros((len(walkers), 3, M), dtype=np.uint32)
    m = recs["m"][walkers].astype(np.int32)
    for b, w in enumerate(walkers):
        words[b] = recs["cur_v"][w][:, :M].astype(np.uint32)
        for k in range(K):
            if pairs[b, k, 0] >= 0:
                rows.append(b)
                ref_idx.append((b, k))
    assert rows, "no candidate pairs"
    tw = words[rows]
    tm = m[rows]
    tp = np.array([pairs[b, k] for b, k in ref_idx], dtype=np.int32)
    s_ref, _ = SR.forward(p.astype(np.float64), layout, tw, tm, tp, nn)
    s_gpu = np.array([scores[b, k] for b, k in ref_idx])
    assert _rel(s_gpu, s_ref) < 1e-2, _rel(s_gpu, s_ref)
    for b, w in enumerate(walkers):
        valid = pairs[b, :, 0] >= 0
        if not valid.any():
            assert forced[w, 0] == -1
            continue
        k = int(np.argmax(np.where(valid, scores[b], -np.inf)))
        assert tuple(forced[w, :3]) == tuple(pairs[b, k]) and forced[w, 3] == 1
        # every candidate is an equal-factor pair of the walker's scheme
        for kk in np.nonzero(valid)[0]:
            c, t1, t2 = pairs[b, kk]
            assert t1 < t2 < m[b] and recs["cur_v"][w][c][t1] == recs["cur_v"][w][c][t2]
