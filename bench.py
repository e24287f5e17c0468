"""Benchmark: flip-graph random walk throughput on one MI355X.

Workload (3x3x3 over Z2): 65536 walkers start from the naive rank-27
scheme, plus-transition after a 20000-flip plateau, target rank 23.  One
step is one launch of the walk kernel: --iterations walk iterations per
walker, followed by the GPU Brent check of every walker's scheme.  Inputs
are resident on the GPU before timing starts.

value = applied flips per second over the timed steps (whole job).
roofline: the walk kernel is bound by VALU integer throughput.  Algorithmic
work per applied flip at rank m is W = 2 + 2m word operations (the two XOR
updates and the equality test of the two new factors against all m terms);
achieved = sum(W) / (sum of walk-kernel durations from HIP events), peak =
CUs * clock * 128 lanes (gfx950 SIMDs execute two wave64 VALU instructions
per 4 cycles, from two waves), checked on the device by fg_valu_probe.
cpu_baseline: the oracle's restatement of the reference walk (history-
ordered FlipSet, as the reference keeps it), built -O3 -march=native on this
host, on every CPU this process may use; median of 5 bounded runs of the same
workload.
ttt: the second half of the metric -- time to a verified 4x4x4 Z2 scheme of
rank <= 49, uniform walks vs learned + uniform proposals (3 seeds per arm in
the default line, 10 with --workload ttt).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--iterations", type=int, default=10000)
    ap.add_argument("--n", type=int, nargs=3, default=[3, 3, 3])
    ap.add_argument("--ring", default="z2", choices=["z2", "zt"], help="z2: GF(2); zt: integers, coefficients in {-1,0,1}")
    ap.add_argument("--plus-iterations", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--word-bits", type=int, default=0, choices=[0, 16, 32],
                    help="Z2 LDS factor words (0: the runtime's choice)")
    ap.add_argument("--partner-layout", type=int, default=0, choices=[0, 1, 2, 3],
                    help="Z2 partner bookkeeping: 0 auto, 1 lists, 2 class masks, 3 class masks with one lane per walker")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttt", action="store_true", help="leave the time-to-target A/B out of the default line")
    ap.add_argument("--ttt-seeds-default", type=int, default=3, help="seeds per arm of the A/B in the default line")
    ap.add_argument("--ttt-train-default", type=int, default=1000, help="training steps of the A/B in the default line")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c3", "c4", "ttt", "propose", "descent"],
                    help="c2: the throughput bench (default); c1: 2x2x2 time to rank 7; c3: batched uneven formats; "
                         "c4: integer-ring walks + additions reducer + complexity minimizer; ttt: learned vs uniform")
    ap.add_argument("--launches", type=int, default=0, help="launch budget of the c1/c3/c4/descent workloads (0: default)")
    ap.add_argument("--expand-probability", type=float, default=0.0, help="--expand-probability of the descent workload")
    tt = ap.add_argument_group("time to target (learned vs uniform proposals, 4x4x4 Z2)")
    tt.add_argument("--ttt", action="store_true", help="run the time-to-target study instead of the throughput bench")
    tt.add_argument("--seeds", type=int, default=10)
    tt.add_argument("--ttt-n", type=int, default=4, help="cube size n of the n x n x n Z2 format (4: rank 64 -> 49, 5: rank 125 -> 100)")
    tt.add_argument("--ttt-walkers", type=int, default=131072)
    tt.add_argument("--ttt-launch", type=int, default=1000, help="walk iterations per launch (one proposal round each)")
    tt.add_argument("--ttt-target", type=int, default=49)
    tt.add_argument("--ttt-timeout", type=float, default=60.0, help="seconds per run before giving up")
    tt.add_argument("--ttt-plus", type=int, default=20000)
    tt.add_argument("--learned-every", type=int, default=2, help="every k-th walker takes the learned proposal")
    tt.add_argument("--candidates", type=int, default=8, help="uniform candidates scored per proposal")
    tt.add_argument("--log-launches", type=int, default=8, help="uniform launches logged for training")
    tt.add_argument("--log-walkers", type=int, default=16384, help="walkers of the logged uniform walk")
    tt.add_argument("--log-warm", type=int, default=30, help="launches of descent before logging starts")
    tt.add_argument("--train-steps", type=int, default=1000)
    tt.add_argument("--train-batch", type=int, default=1024)
    tt.add_argument("--lr", type=float, default=1e-3)
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE readings
    convert the walk kernel's counters into bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, plus_iterations, target_seconds, reps=5):
    """The oracle's restatement of the reference walk (history-ordered FlipSet,
    as the reference keeps it) on every CPU this process may use, one thread
    per CPU with the walkers split between them.  It is built with -O3
    -march=native for this host first (the prebuilt library is the fallback).
    Returns the median flips/s of `reps` runs of about target_seconds / reps
    each, from the naive scheme with the workload's flags."""
    import statistics
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O

    build = "-O3 -march=native, built on this host"
    path = os.path.join(tempfile.mkdtemp(prefix="fg_cpu_baseline_"), "liboracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_OUT={path}"],
                       check=True, capture_output=True, timeout=120)
        O.use_library(path)
    except (subprocess.SubprocessError, OSError, RuntimeError):
        build = "prebuilt oracle/liboracle.so (-O3 -march=x86-64-v2): the native build failed"
    threads = len(os.sched_getaffinity(0))
    walkers = threads * 8

    def run(iters, seed):
        p = O.make_params(iterations=iters, plus_iterations=plus_iterations, select_mode=O.SELECT_REFERENCE)
        recs = O.naive_records(*n, walkers)
        t0 = time.perf_counter()
        recs = O.walk_records(recs, p, seed=seed, nthreads=threads)
        dt = time.perf_counter() - t0
        return int(recs["total_flips"].sum()), dt

    iters = 2000  # calibration: grow until a run takes a measurable time
    while True:
        flips, dt = run(iters, 99)
        if dt >= 1.0 or iters >= 1 << 26:
            break
        iters *= 4
    iters = max(1000, int(iters * (target_seconds / reps) / max(dt, 1e-3)))
    rates, secs = [], []
    for rep in range(reps):
        flips, dt = run(iters, 1 + rep)
        rates.append(flips / dt)
        secs.append(dt)
    return {"value": statistics.median(rates), "unit": "flips/s", "cores": threads, "kind": "port",
            "sample": f"{walkers} walkers x {iters} iterations of the {n[0]}x{n[1]}x{n[2]} Z2 walk from the naive "
                      f"scheme, plus-iterations {plus_iterations}, {reps} runs of {statistics.median(secs):.1f} s "
                      f"(median), {threads} threads",
            "runs": [round(x, 1) for x in rates], "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "threads_available": threads, "build": build}


def profiled_traffic(path):
    """HBM bytes per walk launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (counters cannot be read from inside
    the timed run); None when the summary is absent."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def auc(scores, labels):
    """Area under the ROC curve (Mann-Whitney rank statistic; ties get mean ranks)."""
    import numpy as np
    from scipy.stats import rankdata

    pos = labels > 0.5
    P, N = int(pos.sum()), int((~pos).sum())
    if P == 0 or N == 0:
        return float("nan")
    r = rankdata(scores)
    return float((r[pos].sum() - P * (P + 1) / 2) / (P * N))


def descent_log(seed, W, L, warm, launches, plus, n=4):
    """A uniform n x n x n walk's trajectory log: `warm` launches of descent from
    the naive scheme first, then `launches` logged launches, each walker's first
    flip labelled 1 when its rank falls within that launch."""
    import flipgraphgpu_amd as fg

    e = fg.Engine(W, seed, max_iterations=L, plus_iterations=plus, expand_probability=0.0)
    e.init_naive(n, n, n)
    for _ in range(warm):
        if e.walk(L):
            raise RuntimeError("invalid schemes while logging")
    e.log_enable(max_batches=launches, every=1, horizon=1)
    for _ in range(launches):
        if e.walk(L):
            raise RuntimeError("invalid schemes while logging")
    return e


def heldout_ranking(a, sc, L):
    """How well the trained scorer ranks flips it has not seen: flips logged the
    same way by a fresh uniform walk (seed 2000), scored by the trained network
    and by its untrained initial state (same seed); AUC against the labels."""
    import numpy as np

    import flipgraphgpu_amd as fg

    c, M = a.ttt_n, a.ttt_n ** 3
    e = descent_log(2000, 4096, L, a.log_warm, 4, a.ttt_plus, c)
    n = e.log_info()["rows"]
    words, m, pairs, y = e.log_read(0, n, M)
    e.close()
    sc0 = fg.Scorer(c, c, c, M, seed=7)

    def scores(s):
        out = []
        for b in range(0, n, 16384):
            out.append(s.eval(words[b:b + 16384], m[b:b + 16384], pairs[b:b + 16384])[0])
        return np.concatenate(out)

    res = {"rows": int(n), "positives": int((y > 0.5).sum()), "auc_trained": auc(scores(sc), y),
           "auc_untrained": auc(scores(sc0), y)}
    sc0.close()
    return res


def time_to_target(a, seeds=None, train_steps=None, emit=True):
    """Median wall time over seeds to a verified n x n x n Z2 scheme (n = --ttt-n) of rank <= target,
    uniform walks vs learned + uniform proposals.  The scorer is trained once,
    on trajectories logged by uniform walks of the same shape (seed 1000);
    its training time is reported separately.  A run ends at the first launch
    after which the best scheme has rank <= target and passes the GPU Brent
    check and the library's host re-check.  Returns the result (and prints it
    as one JSON line when emit)."""
    seeds = seeds or a.seeds
    train_steps = train_steps or a.train_steps
    import numpy as np

    import flipgraphgpu_amd as fg

    W, L = a.ttt_walkers, a.ttt_launch
    nc, M = a.ttt_n, a.ttt_n ** 3

    def progress(msg):
        print(msg, file=sys.stderr, flush=True)

    def run(seed, scorer):
        e = fg.Engine(W, seed, max_iterations=L, plus_iterations=a.ttt_plus, expand_probability=0.0)
        e.init_naive(nc, nc, nc)
        t0 = time.perf_counter()
        launches, best = 0, M
        while time.perf_counter() - t0 < a.ttt_timeout:
            if scorer is None:
                bad = e.walk(L)
            else:
                bad = e.walk_learned(scorer, L, candidates=a.candidates, every=a.learned_every, round_=launches)
            if bad:
                raise RuntimeError(f"{bad} invalid schemes")
            launches += 1
            blob, best, _ = e.best()
            if best <= a.ttt_target:
                ok = fg.verify([blob])[0] and fg.verify_host([blob])[0]
                if ok:
                    break
        dt = time.perf_counter() - t0
        e.close()
        return dt, best, launches

    # training data: uniform walks with the trajectory log, from the descent
    # phase (after log_warm launches), labelled by the flip's own launch
    t0 = time.perf_counter()
    e = descent_log(1000, a.log_walkers, L, a.log_warm, a.log_launches, a.ttt_plus, nc)
    info = e.log_info()
    t_log = time.perf_counter() - t0
    progress(f"[ttt] logged {info['rows']} flips ({info['positives']} positive) in {t_log:.1f} s")
    sc = fg.Scorer(nc, nc, nc, M, seed=7)
    t0 = time.perf_counter()
    losses = sc.train(e, steps=train_steps, batch=a.train_batch, lr=a.lr, seed=11)
    t_train = time.perf_counter() - t0
    e.close()
    progress(f"[ttt] trained {train_steps} steps in {t_train:.1f} s, loss {losses[:20].mean():.4f} -> {losses[-50:].mean():.4f}")
    heldout = heldout_ranking(a, sc, L)
    progress(f"[ttt] held-out AUC: trained {heldout['auc_trained']:.4f}, untrained {heldout['auc_untrained']:.4f}")
    res = {"uniform": [], "learned": []}
    for seed in range(1, seeds + 1):
        for arm, scorer in (("uniform", None), ("learned", sc)):
            dt, best, launches = run(seed, scorer)
            res[arm].append({"seed": seed, "seconds": round(dt, 3), "best_rank": int(best), "launches": launches,
                             "reached": bool(best <= a.ttt_target)})
            progress(f"[ttt] seed {seed} {arm}: rank {best} after {launches} launches, {dt:.2f} s")
    sc.close()

    def med(arm):
        v = [r["seconds"] if r["reached"] else float("inf") for r in res[arm]]
        return float(np.median(v))

    mu, ml = med("uniform"), med("learned")
    out = {
        "metric": f"time to verified {nc}x{nc}x{nc} Z2 rank-{a.ttt_target} (median over {seeds} seeds)",
        "value": ml, "unit": "s", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
        "uniform_median_s": mu, "learned_median_s": ml,
        "speedup_learned_vs_uniform": (mu / ml) if ml > 0 else None,
        "config": {"workload": f"{nc}x{nc}x{nc} Z2 from the naive rank-{M} scheme, {W} walkers, {L} iterations per launch, "
                               f"plus-iterations {a.ttt_plus}; learned arm: every {a.learned_every}th walker's first "
                               f"flip of each launch is the scorer's best of {a.candidates} uniform candidates",
                   "walkers": W, "launch_iterations": L},
        "heldout": heldout,
        "training": {"logged_flips": info["rows"], "positives": info["positives"],
                     "log_seconds": round(t_log, 2), "train_seconds": round(t_train, 2),
                     "steps": train_steps, "batch": a.train_batch,
                     "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-50:].mean())},
        "runs": res,
    }
    if emit:
        print(json.dumps(out))
    return out


def propose_flop(n, M, K, D=128, Z=64):
    """Algorithmic FLOPs (2 per multiply-add) of one walker's proposal: three set
    layers over M terms (the first on the F = n1n2 + n2n3 + n3n1 factor bits),
    the two mean products and the pair head of K candidates (2*D inputs, Z
    outputs).  Padding the kernel executes (feature bits rounded up to its
    K-step) is not counted."""
    F = n[0] * n[1] + n[1] * n[2] + n[2] * n[0]
    return 2 * M * (F + 2 * D) * D + 2 * 2 * D * D + 2 * K * 2 * D * Z


def workload_propose(a):
    """Learned-proposal path at 4x4x4 Z2 (rank 64, 131072 walkers): log uniform
    launches, train the scorer for a few steps, then time proposal rounds (every
    2nd walker, 8 candidates) against plain launches.  The fused proposal
    kernel's duration comes from HIP events on the scorer's stream."""
    import numpy as np

    import flipgraphgpu_amd as fg

    W, L, K, every = a.ttt_walkers, a.ttt_launch, a.candidates, a.learned_every
    e = fg.Engine(W, 5, max_iterations=L, plus_iterations=a.ttt_plus, expand_probability=0.0)
    e.init_naive(4, 4, 4)
    e.log_enable(max_batches=12, every=1, horizon=2)
    for q in range(13):
        e.walk(L)
    sc = fg.Scorer(4, 4, 4, 64, seed=7)
    t0 = time.perf_counter()
    steps = 20
    sc.train(e, steps=steps, batch=a.train_batch, lr=a.lr, seed=1)
    t_train = (time.perf_counter() - t0) / steps
    e.close()
    e = fg.Engine(W, 6, max_iterations=L, plus_iterations=a.ttt_plus, expand_probability=0.0)
    e.init_naive(4, 4, 4)
    for q in range(max(a.warmup, 1) + 2):
        e.walk(L)
    sc.propose(e, candidates=K, every=every, round_=0)
    kern, rounds, learned, plain = [], [], [], []
    for q in range(a.steps):
        t0 = time.perf_counter()
        sc.propose(e, candidates=K, every=every, round_=q + 1)
        rounds.append(time.perf_counter() - t0)
        kern.append(sc.last_propose_ms())
        t0 = time.perf_counter()
        e.walk_forced(L)
        learned.append(time.perf_counter() - t0 + rounds[-1])
    for q in range(a.steps):
        t0 = time.perf_counter()
        e.walk(L)
        plain.append(time.perf_counter() - t0)
    sc.close()
    e.close()
    B = (W + every - 1) // every
    flop = propose_flop((4, 4, 4), 64, K) * B
    k_ms = float(np.mean(kern))
    achieved = flop / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "fused proposal round time, 4x4x4 Z2", "value": round(1e3 * float(np.mean(rounds)), 3),
        "unit": "ms", "higher_is_better": False, "n_gpus": 1, "steps": a.steps, "dtype": "bf16",
        "data": "synthetic", "vs_baseline": None,
        "propose_kernel_ms": round(k_ms, 3),
        "learned_launch_ms": round(1e3 * float(np.mean(learned)), 3),
        "plain_launch_ms": round(1e3 * float(np.mean(plain)), 3),
        "train_step_ms": round(1e3 * t_train, 3),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 4), "traffic": None},
        "config": {"workload": f"4x4x4 Z2 rank 64, {W} walkers, every {every}th walker proposes ({B} proposals), "
                               f"{K} candidates each, launches of {L} iterations", "walkers": W, "proposing": B,
                   "candidates": K, "flop_per_round": flop},
    }
    print(json.dumps(out))


def workload_c1(a):
    """2x2x2 Z2 from the naive rank-8 scheme, 1024 walkers, plus after a
    10000-flip plateau: wall time to a verified rank-7 scheme (flip budget 1e6
    per walker)."""
    import flipgraphgpu_amd as fg

    e = fg.Engine(1024, a.seed, max_iterations=1000, plus_iterations=10000, expand_probability=0.0)
    e.init_naive(2, 2, 2)
    c0 = e.counters()
    t0 = time.perf_counter()
    launches, best = 0, 8
    while launches < 1000:
        if e.walk(1000):
            raise RuntimeError("invalid schemes")
        launches += 1
        blob, best, _ = e.best()
        if best <= 7 and fg.verify([blob])[0] and fg.verify_host([blob])[0]:
            break
    dt = time.perf_counter() - t0
    flips = e.counters()["flips"] - c0["flips"]
    e.close()
    print(json.dumps({"metric": "time to verified 2x2x2 Z2 rank 7", "value": dt, "unit": "s",
                      "higher_is_better": False, "n_gpus": 1, "best_rank": int(best), "launches": launches,
                      "flips": int(flips), "flips_per_s": flips / dt,
                      "config": {"workload": "2x2x2 Z2, naive rank-8 start, 1024 walkers, plus-iterations 10000, "
                                             "launches of 1000 iterations (budget 1e6 per walker)"}}))


def workload_c3(a):
    """Five uneven Z2 formats in one persistent pool: 262144 walkers split
    1:2:4:8:16, autotuned grid, per-format best ranks after a launch budget
    against target = naive rank - 20%."""
    import numpy as np

    import flipgraphgpu_amd as fg

    formats = [(2, 2, 3), (2, 3, 4), (3, 3, 4), (3, 4, 5), (4, 4, 4)]
    weights = [1, 2, 4, 8, 16]
    W, L = 262144, 1000
    launches = a.launches or 40
    e = fg.Engine(W, a.seed, max_iterations=L, plus_iterations=20000, expand_probability=0.0)
    e.init_batched(formats, weights)
    grid = e.autotune(200)
    c0 = e.counters()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    for _ in range(launches):
        if e.walk(L):
            raise RuntimeError("invalid schemes")
        kernel_ms += e.last_walk_ms()
    dt = time.perf_counter() - t0
    c1 = e.counters()
    summ = e.summary()
    n = summ["n"]
    per = {}
    for f in formats:
        sel = np.all(n == np.array(f), axis=1)
        naive = f[0] * f[1] * f[2]
        best = int(summ["best_m"][sel].min())
        per["x".join(map(str, f))] = {"walkers": int(sel.sum()), "naive": naive, "target": round(naive * 0.8, 1),
                                     "best": best, "reached": best <= naive * 0.8}
    e.close()
    flips = c1["flips"] - c0["flips"]
    print(json.dumps({"metric": "flips/s (batched uneven Z2 formats, 262144 walkers)", "value": flips / dt,
                      "unit": "flips/s", "n_gpus": 1, "launches": launches, "seconds": dt,
                      "kernel_ms_per_launch": kernel_ms / launches, "autotuned_grid": grid, "formats": per,
                      "config": {"workload": "Z2 formats 2x2x3/2x3x4/3x3x4/3x4x5/4x4x4 from naive, walkers 1:2:4:8:16 "
                                             "of 262144 (largest remainder), plus-iterations 20000, launches of 1000 "
                                             "iterations, one persistent launch for all formats"}}))


def workload_c4(a):
    """Integer-ring ({-1,0,1}) walks of 3x3x3 and 4x4x4, 65536 walkers each,
    from the naive scheme; the best scheme found goes to the additions
    reducer and the walkers' schemes of that rank to the complexity
    minimizer.  Reports additions before and after."""
    import numpy as np

    import flipgraphgpu_amd as fg

    out = {"metric": "additions of integer-ring schemes before / after reduction", "n_gpus": 1, "formats": {}}
    for n, launches in (((3, 3, 3), a.launches or 30), ((4, 4, 4), a.launches or 30)):
        W = 65536
        e = fg.Engine(W, a.seed, ring=fg.RING_ZT, max_iterations=1000, plus_iterations=20000,
                      expand_probability=0.0)
        e.init_naive(*n)
        c0 = e.counters()
        t0 = time.perf_counter()
        for _ in range(launches):
            if e.walk(1000):
                raise RuntimeError("invalid schemes")
        dt = time.perf_counter() - t0
        flips = e.counters()["flips"] - c0["flips"]
        blob, rank, _ = e.best()
        assert fg.verify([blob], ring=fg.RING_ZT)[0] and fg.verify_host([blob], fg.RING_ZT)[0]
        recs = e.read_walkers()
        cands = [w for w in range(W) if recs[w]["best_m"] == rank][:256]
        blobs = [e.walker_scheme(w, best=True) for w in cands]
        e.close()
        with fg.Engine(1, 1, ring=fg.RING_ZT) as ne:
            ne.init_naive(*n)
            naive_blob = ne.walker_scheme(0, best=False)
        naive_cx = int(fg.complexity_host([naive_blob], fg.RING_ZT)[0])
        cx_found = int(fg.complexity_host(blobs, fg.RING_ZT).min())
        # complexity minimizer over the walkers' best schemes of that rank
        t1 = time.perf_counter()
        z = fg.Engine(4096, a.seed, ring=fg.RING_ZT, max_iterations=10000, check_reduce=False)
        z.init_minimizer(np.concatenate([np.array([len(blobs)], dtype=np.int32)] + blobs))
        best_cx = None
        for it in range(8):
            if z.minimize():
                raise RuntimeError("invalid schemes in the minimizer")
            bc, _ = z.complexities()
            bw = int(np.argmin(bc))
            if best_cx is not None and bc[bw] >= best_cx:
                break
            best_cx = int(bc[bw])
            z.broadcast_best(bw, list(range(0, 4096, 2)))
        mblob, mcx, _ = z.best()
        z.close()
        t_min = time.perf_counter() - t1
        # additions reducer on the minimized scheme
        t2 = time.perf_counter()
        r = fg.Reducer(mblob, count=512, seed=a.seed)
        no_imp = 0
        while no_imp < 3:
            r.run()
            no_imp = 0 if r.update() else no_imp + 1
        b = r.best()
        r.close()
        t_red = time.perf_counter() - t2
        out["formats"]["x".join(map(str, n))] = {
            "walkers": W, "launches": launches, "walk_seconds": round(dt, 2), "flips_per_s": flips / dt,
            "best_rank": int(rank), "naive_rank": n[0] * n[1] * n[2],
            "additions_naive_scheme": int(naive_cx),
            "additions_found_scheme": int(cx_found),
            "additions_after_minimizer": int(mcx), "minimizer_seconds": round(t_min, 2),
            "reducer_naive_additions": int(sum(b["naive"])), "reducer_reduced_additions": int(b["total"]),
            "reducer_fresh_variables": int(b["total_fresh"]), "reducer_seconds": round(t_red, 2),
        }
        print(f"[c4] {n}: rank {rank}, additions {cx_found} -> {mcx} (minimizer) -> {b['total']} (reducer)",
              file=sys.stderr, flush=True)
    out["config"] = {"workload": "integer ring {-1,0,1}, 65536 walkers per format from naive, launches of 1000 "
                                 "iterations; best rank's schemes -> complexity minimizer (4096 walkers) -> additions "
                                 "reducer (512 reducers, until 3 passes without improvement)"}
    print(json.dumps(out))


def workload_descent(a):
    """Rank descent of 4x4x4 walks from the naive rank-64 scheme over both
    rings (65536 walkers, launches of 1000 iterations, --plus-iterations and
    --expand-probability as given): best and mean rank every 20 launches."""
    import numpy as np

    import flipgraphgpu_amd as fg

    launches = a.launches or 200
    out = {"metric": "rank descent of 4x4x4 walks from the naive scheme", "n_gpus": 1, "rings": {}}
    for name, ring in (("z2", fg.RING_Z2), ("zt", fg.RING_ZT)):
        e = fg.Engine(65536, a.seed, ring=ring, max_iterations=1000, plus_iterations=a.plus_iterations,
                      expand_probability=a.expand_probability)
        e.init_naive(4, 4, 4)
        c0 = e.counters()
        cw = c0
        t0 = time.perf_counter()
        traj = []
        for q in range(launches):
            if e.walk(1000):
                raise RuntimeError("invalid schemes")
            if (q + 1) % 20 == 0:
                s = e.summary()
                c = e.counters()
                # share of iterations whose tryFlip found a flip (the rest had
                # no admissible candidate and took the expand fallback)
                rate = (c["flips"] - cw["flips"]) / max(1, c["iterations"] - cw["iterations"])
                cw = c
                ranks, counts = np.unique(s["m"], return_counts=True)
                traj.append({"launches": q + 1, "best_rank": int(s["best_m"].min()),
                             "mean_rank": round(float(np.mean(s["m"])), 4), "flip_success": round(float(rate), 4),
                             "rank_histogram": {int(r): int(k) for r, k in zip(ranks, counts) if k >= 64}})
                print(f"[descent] {name} {traj[-1]}", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
        out["rings"][name] = {"seconds": round(dt, 2), "flips_per_s": (e.counters()["flips"] - c0["flips"]) / dt,
                              "trajectory": traj}
        e.close()
    out["config"] = {"workload": f"4x4x4, 65536 walkers, launches of 1000 iterations, plus-iterations "
                                 f"{a.plus_iterations}, expand-probability {a.expand_probability}"}
    print(json.dumps(out))


def main():
    a = parse()
    if a.workload in ("c1", "c3", "c4", "descent") and a.gpus == 1:
        {"c1": workload_c1, "c3": workload_c3, "c4": workload_c4, "descent": workload_descent}[a.workload](a)
        return
    if a.workload == "propose" and a.gpus == 1:
        workload_propose(a)
        return
    if a.workload == "ttt":
        a.ttt = True
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.ttt:
        if a.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
            sys.exit(2)
        time_to_target(a)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import flipgraphgpu_amd as fg

    info = fg.device_info()
    n = tuple(a.n)
    zt = a.ring == "zt"
    e = fg.Engine(a.walkers, a.seed, ring=fg.RING_ZT if zt else fg.RING_Z2, max_iterations=a.iterations,
                  plus_iterations=a.plus_iterations, expand_probability=0.0, word_bits=0 if zt else a.word_bits,
                  partner_layout=0 if zt else a.partner_layout)
    e.init_naive(*n)
    for _ in range(a.warmup):
        bad = e.walk(a.iterations)
        if bad:
            raise RuntimeError(f"{bad} invalid schemes after warmup")
    c0 = e.counters()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        bad = e.walk(a.iterations)
        kernel_ms += e.last_walk_ms()
        if bad:
            raise RuntimeError(f"{bad} invalid schemes")
    dt = time.perf_counter() - t0
    c1 = e.counters()
    flips = c1["flips"] - c0["flips"]
    sum_m = c1["sum_m"] - c0["sum_m"]
    # W = 2 + 2m word operations per Z2 flip, 4 + 4m for the integer ring (two words per factor)
    ops = (4 if zt else 2) * (flips + sum_m)
    recs = e.read_walkers()
    hist = {int(k): int(v) for k, v in zip(*__import__("numpy").unique(recs["best_m"], return_counts=True))}
    _, best_rank, _ = e.best()
    e.close()

    # gfx950 VALU peak: a SIMD executes two wave64 instructions per 4 cycles
    # (from two waves), 128 lanes per CU-clock, the rate behind the published
    # FP32 vector figure; fg_valu_probe measures it on this device
    peak = info["cu"] * info["clock_khz"] * 1e3 * 128
    probe_ops, probe_ms = fg.valu_probe()
    traffic_src = "profiles/walk_z2_3x3x3_hbm_traffic.md"
    default_shape = n == (3, 3, 3) and a.walkers == 65536 and a.iterations == 10000 and not zt
    traffic = profiled_traffic(os.path.join(ROOT, traffic_src)) if default_shape else None
    achieved = ops / (kernel_ms / 1e3) if kernel_ms > 0 else 0.0
    out = {
        "metric": ("flips/s; time to verified 4x4x4 Z2 rank-49, learned vs uniform flips" if default_shape else
                   f"flips/s ({'integer-ring' if zt else 'Z2'} flip-graph random walk, {n[0]}x{n[1]}x{n[2]}, "
                   f"{a.walkers} walkers)"),
        "value": flips / dt,
        "unit": "flips/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"{n[0]}x{n[1]}x{n[2]} {'integer-ring' if zt else 'Z2'} walk from the naive rank-{n[0] * n[1] * n[2]} scheme, "
                        f"{a.walkers} walkers, {a.iterations} iterations per walker per step, "
                        f"plus-iterations {a.plus_iterations}, Brent check of every scheme after each step",
            "walkers": a.walkers,
            "iterations_per_step": a.iterations,
            "best_rank": best_rank,
            "best_rank_histogram": hist,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TOP/s",
            "frac": achieved / peak if peak else None,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "traffic_source": traffic_src if traffic is not None else None,
            "kernel_ms_per_step": kernel_ms / a.steps,
            "ops_per_flip": ops / flips if flips else None,
            "peak_measured": probe_ops / 1e12,
            "peak_measured_by": f"fg_valu_probe: independent v_xor_b32 chains, 16384 x 256 threads, {probe_ms:.3f} ms",
            "peak_one_wave_per_simd": peak / 2e12,
            "device": info,
        },
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, a.plus_iterations, a.cpu_seconds)
    if not a.no_ttt:
        t = time_to_target(a, seeds=a.ttt_seeds_default, train_steps=a.ttt_train_default, emit=False)
        out["ttt"] = {"metric": t["metric"], "uniform_median_s": t["uniform_median_s"],
                      "learned_median_s": t["learned_median_s"], "speedup": t["speedup_learned_vs_uniform"],
                      "config": t["config"]["workload"], "heldout": t["heldout"], "training": t["training"],
                      "runs": t["runs"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
