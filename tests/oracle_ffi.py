"""ctypes access to the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

RING_Z2 = 0
RING_ZT = 1
SELECT_CANONICAL = 0
SELECT_REFERENCE = 1
REC_MAX_M = 256

RECORD_DTYPE = np.dtype([
    ("n", np.int32, (3,)),
    ("m", np.int32),
    ("best_m", np.int32),
    ("best_rank", np.int32),
    ("flips_count", np.int32),
    ("aux_idx", np.int32),
    ("it", np.uint64),
    ("aux", np.uint64),
    ("total_flips", np.uint64),
    ("sum_m", np.uint64),
    ("cur_v", np.uint64, (3, REC_MAX_M)),
    ("cur_s", np.uint64, (3, REC_MAX_M)),
    ("best_v", np.uint64, (3, REC_MAX_M)),
    ("best_s", np.uint64, (3, REC_MAX_M)),
    ("best_n", np.int32, (3,)),
    ("reserved", np.int32),
])


class WalkParams(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("random_iterations", ctypes.c_int),
        ("plus_iterations", ctypes.c_int),
        ("use_reduce", ctypes.c_int),
        ("thr_reduce", ctypes.c_uint32),
        ("use_expand", ctypes.c_int),
        ("thr_expand", ctypes.c_uint32),
        ("thr_best_eq", ctypes.c_uint32),
        ("select_mode", ctypes.c_int),
        ("check_reduce", ctypes.c_int),
        ("zt_visit_cap", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("use_sandwich", ctypes.c_int),
        ("thr_sandwich", ctypes.c_uint32),
        ("use_basis", ctypes.c_int),
        ("thr_basis", ctypes.c_uint32),
        ("min_window", ctypes.c_int),
        ("look_period", ctypes.c_int),
        ("look_candidates", ctypes.c_int),
    ]


def threshold(p, max_iterations):
    """P(u * maxIterations < p) for u uniform in (0, 1] as a 32-bit threshold."""
    x = int(p / max_iterations * 4294967296.0)
    return max(0, min(x, 0xFFFFFFFF))


THR_BEST_EQ = 42949673  # 0.01 * 2^32, rounded up


def make_params(iterations=1000, max_iterations=None, plus_iterations=8000, p_expand=0.0,
                p_reduce=0.0, random_iterations=False, select_mode=SELECT_CANONICAL,
                check_reduce=True, zt_visit_cap=16, mode=0, p_sandwich=0.0, p_basis=0.0, min_window=0,
                look_period=0, look_candidates=0):
    mi = max_iterations if max_iterations is not None else iterations
    return WalkParams(iterations, mi, int(random_iterations), plus_iterations,
                      int(p_reduce > 0), threshold(p_reduce, mi),
                      int(p_expand > 0), threshold(p_expand, mi),
                      THR_BEST_EQ, select_mode, int(check_reduce), zt_visit_cap, mode,
                      int(p_sandwich > 0), threshold(p_sandwich, mi), int(p_basis > 0), threshold(p_basis, mi),
                      min_window, look_period, look_candidates)


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def use_library(path):
    """Load the oracle from `path` (a build of the same sources, e.g. with
    -march=native for a timing) instead of oracle/liboracle.so; call before
    the first lib()."""
    global LIB_PATH, _lib
    if _lib is not None and LIB_PATH != path:
        raise RuntimeError("the oracle library is already loaded")
    LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            _build()
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.fgo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.fgo_blob_validate.argtypes = [ctypes.c_int, i32p]
        L.fgo_blob_complexity.argtypes = [ctypes.c_int, i32p]
        L.fgo_blob_naive.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        L.fgo_blob_replay.argtypes = [ctypes.c_int, i32p, i32p, ctypes.c_int, i32p]
        L.fgo_blob_pairs.argtypes = [ctypes.c_int, i32p, ctypes.c_int, u32p, ctypes.c_int]
        L.fgo_walk_records.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.POINTER(WalkParams), ctypes.c_int, ctypes.c_int]
        L.fgo_naive_records.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int]
        L.fgo_record_validate.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().fgo_philox4x32_10(c, k, o)
    return list(o)


def naive_blob(n1, n2, n3, ring=RING_Z2):
    m = n1 * n2 * n3
    out = np.zeros(4 + m * (n1 * n2 + n2 * n3 + n3 * n1), dtype=np.int32)
    _, p = _i32(out)
    n = lib().fgo_blob_naive(ring, n1, n2, n3, p)
    return out[:n]


def validate(blob, ring=RING_Z2):
    b, p = _i32(blob)
    return lib().fgo_blob_validate(ring, p)


def complexity(blob, ring=RING_Z2):
    b, p = _i32(blob)
    return lib().fgo_blob_complexity(ring, p)


def replay(blob, ops, ring=RING_Z2, max_m=350):
    b, p = _i32(blob)
    o = np.zeros((max(len(ops), 1), 8), dtype=np.int32)
    for q, op in enumerate(ops):
        o[q, :len(op)] = op
    o, po = _i32(o)
    n1, n2, n3 = int(blob[0]), int(blob[1]), int(blob[2])
    out = np.zeros(4 + max_m * (n1 * n2 + n2 * n3 + n3 * n1), dtype=np.int32)
    _, pout = _i32(out)
    n = lib().fgo_blob_replay(ring, p, po, len(ops), pout)
    assert n > 0
    return out[:n]


def pairs(blob, ring=RING_Z2, mode=0):
    b, p = _i32(blob)
    cap = 200000
    out = np.zeros(cap, dtype=np.uint32)
    n = lib().fgo_blob_pairs(ring, p, mode, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
    return [(int(x) >> 24, (int(x) >> 12) & 0xFFF, int(x) & 0xFFF) for x in out[:n]]


def naive_records(n1, n2, n3, count, ring=RING_Z2):
    recs = np.zeros(count, dtype=RECORD_DTYPE)
    rc = lib().fgo_naive_records(ring, n1, n2, n3, recs.ctypes.data, count)
    assert rc == 0
    return recs


def walk_records(recs, params, seed, ring=RING_Z2, walker_base=0, launches=1, nthreads=1):
    recs = np.array(recs, dtype=RECORD_DTYPE, copy=True, order="C")
    rc = lib().fgo_walk_records(ring, recs.ctypes.data, len(recs), walker_base, seed,
                                ctypes.byref(params), launches, nthreads)
    assert rc == 0
    return recs


def walk_records_forced(recs, params, seed, forced=None, walker_base=0):
    """One Z2 launch with proposed first flips; returns (records, log [n][4], minm [n])."""
    recs = np.array(recs, dtype=RECORD_DTYPE, copy=True, order="C")
    L = lib()
    L.fgo_walk_records_forced.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                          ctypes.POINTER(WalkParams), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
    fz = None if forced is None else np.ascontiguousarray(forced, dtype=np.int32)
    log = np.zeros((len(recs), 4), dtype=np.int32)
    mm = np.zeros(len(recs), dtype=np.int32)
    rc = L.fgo_walk_records_forced(recs.ctypes.data, len(recs), walker_base, seed, ctypes.byref(params),
                                   None if fz is None else fz.ctypes.data, log.ctypes.data, mm.ctypes.data)
    assert rc == 0
    return recs, log, mm


def resize_records(recs, seed, thr_resize, max_rank, max_coef=None, walker_base=0, ring=RING_Z2):
    """One resizeKernel pass over walker records (max_coef: 32 for Z2, 16 for ZT)."""
    if max_coef is None:
        max_coef = 16 if ring == RING_ZT else 32
    recs = np.array(recs, dtype=RECORD_DTYPE, copy=True, order="C")
    L = lib()
    L.fgo_resize_records.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    rc = L.fgo_resize_records(ring, recs.ctypes.data, len(recs), walker_base, seed, thr_resize, max_rank, max_coef)
    assert rc == 0
    return recs


RESIZE_OPS = {"project": 0, "extend": 1, "product": 2, "merge": 3, "product_scheme": 4, "swap_size": 5,
              "swap_rows": 6, "swap_columns": 7}


def resize_op(blob, op, a=0, b=0, other=None, ring=RING_Z2):
    """One resize or basis operator of the oracle on a blob."""
    L = lib()
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.fgo_blob_resize_op.argtypes = [ctypes.c_int, i32p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
    x, px = _i32(blob)
    po = None
    if other is not None:
        o, po = _i32(other)
    out = np.zeros(4 + 350 * 3 * 64, dtype=np.int32)
    n = L.fgo_blob_resize_op(ring, px, po, RESIZE_OPS[op], a, b, out.ctypes.data_as(i32p))
    assert n > 0
    return out[:n].copy()


def record_valid(rec_array, index, best=False, ring=RING_Z2):
    return lib().fgo_record_validate(ring, rec_array[index:index + 1].ctypes.data, int(best))


# ---- scheme blobs <-> python lists -----------------------------------
def blob_to_terms(blob):
    n1, n2, n3, m = (int(x) for x in blob[:4])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    p = 4
    comps = []
    for c in range(3):
        rows = []
        for _ in range(m):
            rows.append([int(x) for x in blob[p:p + nn[c]]])
            p += nn[c]
        comps.append(rows)
    return (n1, n2, n3), comps


def terms_to_blob(n, comps):
    n1, n2, n3 = n
    m = len(comps[0])
    out = [n1, n2, n3, m]
    for c in range(3):
        for row in comps[c]:
            out.extend(row)
    return np.array(out, dtype=np.int32)
