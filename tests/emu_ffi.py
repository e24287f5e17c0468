"""ctypes access to the host emulation of the walk kernels (tests/emu)."""
import ctypes
import os
import subprocess

import numpy as np

import oracle_ffi as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "emu", "libemu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu")], check=True)
        L = ctypes.CDLL(LIB)
        L.emu_walk_records.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64] + \
            [ctypes.c_int] * 4 + [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


LAYOUTS = {"auto": 0, "wide": 1, "half": 2, "masks": 3, "solo": 4}


def walk(recs, params, seed, launches=1, ring=O.RING_Z2, forced=None, learn=False, wide=False, resize=None,
         layout="auto", trace_cap=0):
    """Emulated launches; with learn (Z2) the first launch takes the proposed
    pairs `forced` [walkers][4] and the result adds (log_meta, minm).  Z2
    layouts as the runtime picks them (auto: class masks up to 32 terms and
    16 coefficients, else 16-bit words with partner lists up to 16
    coefficients, else 32-bit words), or forced: "wide", "half", "masks",
    "solo" (class masks, one lane per walker; wide=True is "wide").  resize = (thr_resize, max_rank, max_coef) runs the
    resize mode instead of the walk.  trace_cap > 0 runs the trace mode (one
    launch) and the result adds (ops [walkers][trace_cap][8], counts)."""
    if wide:
        layout = "wide"
    recs = np.ascontiguousarray(recs.copy())
    err = ctypes.c_int32()
    p = params
    mode = 2 if learn else (4 if resize is not None else p.mode)
    if mode == 0 and (p.use_sandwich or p.use_basis):
        mode = 5  # the meta-walk kernel
    if mode == 0 and p.look_candidates > 0:
        mode = 6  # the lookahead kernel
    if trace_cap > 0:
        assert mode == 0 and launches == 1, "traces are recorded for one plain launch"
        mode = 7
    tr = np.zeros((len(recs), max(trace_cap, 1), 8), dtype=np.int32)
    tn = np.zeros(len(recs), dtype=np.int32)
    thr, rmax, cmax = resize if resize is not None else (0, 0, 0)
    fz = None if forced is None else np.ascontiguousarray(forced, dtype=np.int32)
    lm = np.zeros((len(recs), 4), dtype=np.int32)
    mm = np.zeros(len(recs), dtype=np.int32)
    rc = lib().emu_walk_records(ring, recs.ctypes.data, len(recs), seed, p.iterations, p.max_iterations,
                                p.random_iterations, p.plus_iterations, p.use_reduce, p.thr_reduce,
                                p.use_expand, p.thr_expand, p.thr_best_eq, p.check_reduce, launches,
                                p.zt_visit_cap, mode, None if fz is None else fz.ctypes.data,
                                lm.ctypes.data, mm.ctypes.data, LAYOUTS[layout], thr, rmax, cmax,
                                p.use_sandwich, p.thr_sandwich, p.use_basis, p.thr_basis, p.look_period,
                                p.look_candidates, tr.ctypes.data if trace_cap > 0 else None,
                                tn.ctypes.data if trace_cap > 0 else None, trace_cap, ctypes.byref(err))
    assert rc == 0
    if trace_cap > 0:
        return recs, err.value, tr[:, :trace_cap], tn
    if learn:
        return recs, err.value, lm, mm
    return recs, err.value
