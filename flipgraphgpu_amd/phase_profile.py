"""Per-phase cycle profile of the Z2 walk kernel.

Loads the phase profile build of the library (make -C flipgraphgpu_amd
phases: the same kernels with s_memtime marks between the phases of an
iteration) and reports, for each workload, the share of wave cycles spent in
each phase and the cycles per walk iteration.  Marks wait for outstanding LDS
operations, so this build runs a few percent slower than the product; the
shares, not the absolute times, are what it is for.

Usage: python -m flipgraphgpu_amd.phase_profile [--layout 0|1|2|3] [--word-bits 0|16|32]
"""
import argparse
import ctypes
import json
import os

from . import _lib

PHASES = ["rng", "select", "scan", "classes", "exchange", "leave_join", "bookkeeping_copy", "loop_tail",
          "reduction", "zero_factor"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs=3, default=[3, 3, 3])
    ap.add_argument("--walkers", type=int, default=65536)
    ap.add_argument("--iterations", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--layout", type=int, default=0,
                    help="partner_layout: 0 auto, 1 lists, 2 class masks, 3 class masks with one lane per walker")
    ap.add_argument("--word-bits", type=int, default=0)
    a = ap.parse_args()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_phase", "libflipgraph.so")
    if not os.path.exists(path):
        raise SystemExit(f"{path} is missing: build it with make -C flipgraphgpu_amd phases")
    _lib.LIB_PATH = path
    L = _lib.load_library()
    L.fg_phase_read.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    L.fg_phase_read_solo.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    e = _lib.Engine(a.walkers, 1, max_iterations=a.iterations, plus_iterations=20000, word_bits=a.word_bits,
                    partner_layout=a.layout)
    e.init_naive(*a.n)
    for _ in range(a.warmup):
        e.walk(a.iterations)
    L.fg_phase_reset()
    L.fg_phase_reset_solo()
    c0 = e.counters()
    e.walk(a.iterations)
    ms = e.last_walk_ms()
    c1 = e.counters()
    layout, _ = e.layout_info()
    e.close()
    buf = (ctypes.c_ulonglong * 16)()
    L.fg_phase_read(buf)
    solo = (ctypes.c_ulonglong * 16)()
    L.fg_phase_read_solo(solo)  # the one-lane layout's kernels keep their own sums
    for q in range(16):
        buf[q] += solo[q]
    total = sum(buf[:10])
    iters = buf[15]
    print(json.dumps({
        "workload": f"{a.n[0]}x{a.n[1]}x{a.n[2]} Z2, {a.walkers} walkers, one launch of {a.iterations} iterations "
                    f"after {a.warmup} warm-up launch(es)",
        "layout": _lib.Engine.LAYOUTS[layout], "kernel_ms": round(ms, 3),
        "flips_per_s": (c1["flips"] - c0["flips"]) / ms * 1e3,
        "cycles_per_iteration_per_lane": round(total / max(iters, 1)),
        "phase_share": {PHASES[q]: round(buf[q] / total, 4) for q in range(10)}}))


if __name__ == "__main__":
    main()
