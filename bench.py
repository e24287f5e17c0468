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
CUs * clock * 64 lanes.
cpu_baseline: the oracle's restatement of the reference walk (history-
ordered FlipSet, as the reference keeps it) on the host cores, on a bounded
sample of the same workload.
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
    ap.add_argument("--plus-iterations", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration) and exit")
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


def cpu_baseline(n, plus_iterations, target_seconds):
    """Oracle (reference FlipSet order) on the host cores; returns flips/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O

    threads = max(1, min(16, os.cpu_count() or 1))
    walkers = threads * 8
    iters = 1000
    while True:
        p = O.make_params(iterations=iters, plus_iterations=plus_iterations, select_mode=O.SELECT_REFERENCE)
        recs = O.naive_records(*n, walkers)
        t0 = time.perf_counter()
        recs = O.walk_records(recs, p, seed=1, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= target_seconds * 0.5 or iters >= 1 << 26:
            break
        iters = int(iters * min(16.0, max(2.0, target_seconds / max(dt, 1e-3))))
    flips = int(recs["total_flips"].sum())
    return {"value": flips / dt, "unit": "flips/s", "cores": threads, "kind": "port",
            "sample": f"{walkers} walkers x {iters} iterations of the {n[0]}x{n[1]}x{n[2]} Z2 walk "
                      f"from the naive scheme, {dt:.1f} s wall"}


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


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import flipgraphgpu_amd as fg

    info = fg.device_info()
    n = tuple(a.n)
    e = fg.Engine(a.walkers, a.seed, max_iterations=a.iterations, plus_iterations=a.plus_iterations,
                  expand_probability=0.0)
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
    ops = 2 * flips + 2 * sum_m
    recs = e.read_walkers()
    hist = {int(k): int(v) for k, v in zip(*__import__("numpy").unique(recs["best_m"], return_counts=True))}
    _, best_rank, _ = e.best()
    e.close()

    peak = info["cu"] * info["clock_khz"] * 1e3 * 64
    traffic_src = "profiles/walk_z2_3x3x3_hbm_traffic.md"
    default_shape = n == (3, 3, 3) and a.walkers == 65536 and a.iterations == 10000
    traffic = profiled_traffic(os.path.join(ROOT, traffic_src)) if default_shape else None
    achieved = ops / (kernel_ms / 1e3) if kernel_ms > 0 else 0.0
    out = {
        "metric": "flips/s (Z2 flip-graph random walk, 3x3x3, 65536 walkers)",
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
            "workload": f"{n[0]}x{n[1]}x{n[2]} Z2 walk from the naive rank-{n[0] * n[1] * n[2]} scheme, "
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
            "device": info,
        },
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, a.plus_iterations, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
