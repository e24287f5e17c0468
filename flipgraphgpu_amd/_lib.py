"""ctypes binding of libflipgraph.so (include/fg.h).

The shared library holds the HIP kernels and the host runtime; this module
only moves buffers across the C ABI.  It never falls back to CPU code: if
the library is missing, importing the engine raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libflipgraph.so")

RING_Z2 = 0
RING_ZT = 1
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


class FgConfig(ctypes.Structure):
    _fields_ = [
        ("ring", ctypes.c_int32),
        ("walkers", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("plus_iterations", ctypes.c_int32),
        ("random_iterations", ctypes.c_int32),
        ("check_reduce", ctypes.c_int32),
        ("expand_probability", ctypes.c_double),
        ("reduce_probability", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("max_rank", ctypes.c_int32),
        ("zt_visit_cap", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("word_bits", ctypes.c_int32),
        ("sandwiching_probability", ctypes.c_double),
        ("basis_probability", ctypes.c_double),
        ("partner_layout", ctypes.c_int32),
        ("lookahead_period", ctypes.c_int32),
        ("lookahead_candidates", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
    ]


class FgError(RuntimeError):
    pass


_lib = None


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FgError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C flipgraphgpu_amd)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)
    L.fg_last_error.restype = ctypes.c_char_p
    L.fg_create.argtypes = [ctypes.POINTER(FgConfig), ctypes.POINTER(vp)]
    L.fg_destroy.argtypes = [vp]
    L.fg_init_naive.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.fg_init_from_schemes.argtypes = [vp, i32p, ctypes.c_int64]
    L.fg_init_batched.argtypes = [vp, ctypes.c_int, i32p, i32p]
    L.fg_init_minimizer.argtypes = [vp, i32p, ctypes.c_int64]
    L.fg_minimize.argtypes = [vp]
    L.fg_init_flips.argtypes = [vp, i32p, ctypes.c_int64]
    L.fg_resize.argtypes = [vp, ctypes.c_double]
    L.fg_snapshot_best.argtypes = [vp, i32p, ctypes.c_int]
    L.fg_set_best_ranks.argtypes = [vp, i32p]
    L.fg_flip_schemes.argtypes = [vp, ctypes.c_int]
    L.fg_complexities.argtypes = [vp, i32p, i32p]
    L.fg_broadcast_best.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int]
    L.fg_autotune.argtypes = [vp, ctypes.c_int, i32p]
    L.fg_layout_info.argtypes = [vp, i32p, i32p, ctypes.c_int]
    L.fg_walk.argtypes = [vp, ctypes.c_int]
    L.fg_best.argtypes = [vp, i32p, ctypes.c_int64, i32p, i32p]
    L.fg_walker_scheme.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int64]
    L.fg_verify.argtypes = [ctypes.c_int, i32p, ctypes.c_int64, i32p]
    L.fg_verify_noalloc.argtypes = [ctypes.c_int, i32p, ctypes.c_int64, i32p]
    L.fg_verify_host.argtypes = [ctypes.c_int, i32p, ctypes.c_int64, i32p]
    L.fg_complexity_host.argtypes = [ctypes.c_int, i32p, ctypes.c_int64, i32p]
    L.fg_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_uint64)]
    L.fg_last_walk_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.fg_read_walkers.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.fg_summary.argtypes = [vp, i32p, i32p, i32p, i32p, i32p]
    L.fg_write_walkers.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.fg_log_enable.argtypes = [vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    L.fg_log_window.argtypes = [vp, ctypes.c_int]
    L.fg_log_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                              ctypes.POINTER(ctypes.c_int64)]
    L.fg_log_read.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, vp, vp, vp, vp]
    L.fg_set_forced.argtypes = [vp, vp]
    L.fg_get_forced.argtypes = [vp, vp]
    L.fg_walk_forced.argtypes = [vp, ctypes.c_int]
    L.fg_walk_trace.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    L.fg_scorer_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.POINTER(vp)]
    L.fg_scorer_destroy.argtypes = [vp]
    L.fg_scorer_num_params.argtypes = [vp]
    L.fg_scorer_num_params.restype = ctypes.c_int64
    L.fg_scorer_layout.argtypes = [vp, i32p]
    L.fg_scorer_get_params.argtypes = [vp, vp]
    L.fg_scorer_set_params.argtypes = [vp, vp]
    L.fg_scorer_eval.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.fg_scorer_train.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, vp]
    L.fg_scorer_propose.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
    L.fg_walk_learned.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
    L.fg_reducer_create.argtypes = [i32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.POINTER(vp)]
    L.fg_reducer_destroy.argtypes = [vp]
    L.fg_reducer_info.argtypes = [vp, i32p]
    L.fg_reducer_scheme_stats.argtypes = [vp, i32p]
    L.fg_reducer_run.argtypes = [vp, ctypes.c_int]
    L.fg_reducer_update.argtypes = [vp, i32p]
    L.fg_reducer_best.argtypes = [vp, i32p]
    L.fg_reducer_rows.argtypes = [vp, ctypes.c_int, i32p]
    L.fg_reducer_best_part.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int64]
    L.fg_reducer_best_part.restype = ctypes.c_int64
    L.fg_reducer_part.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int64]
    L.fg_reducer_part.restype = ctypes.c_int64
    L.fg_reducer_words.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.fg_device_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    _lib = L
    return L


def _check(rc, what):
    if rc < 0:
        raise FgError(f"{what}: {load_library().fg_last_error().decode()} (rc={rc})")
    return rc


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def device_info():
    L = load_library()
    cu, clk, lds = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(64)
    _check(L.fg_device_info(ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(lds), name, 64), "fg_device_info")
    return {"cu": cu.value, "clock_khz": clk.value, "lds_per_block": lds.value, "arch": name.value.decode()}


def valu_probe(blocks=16384, trips=100):
    """Measured 32-bit integer VALU rate of the current GPU (fg_valu_probe):
    (lane operations per second, kernel ms of the fastest timed launch)."""
    L = load_library()
    ops, ms = ctypes.c_double(), ctypes.c_double()
    _check(L.fg_valu_probe(int(blocks), int(trips), ctypes.byref(ops), ctypes.byref(ms)), "fg_valu_probe")
    return ops.value, ms.value


class Engine:
    """One walker pool on the current GPU (fg_ctx)."""

    def __init__(self, walkers, seed, ring=RING_Z2, max_iterations=10000, plus_iterations=8000,
                 random_iterations=False, check_reduce=True, expand_probability=0.0,
                 reduce_probability=0.0, max_rank=0, zt_visit_cap=16, use_graph=False, grid=0, word_bits=0,
                 sandwiching_probability=0.0, basis_probability=0.0, partner_layout=0, lookahead_period=0,
                 lookahead_candidates=0):
        self.L = load_library()
        self.ring = ring
        self.cfg = FgConfig(ring, walkers, max_iterations, plus_iterations, int(random_iterations),
                            int(check_reduce), expand_probability, reduce_probability, seed, max_rank,
                            zt_visit_cap, int(use_graph), grid, word_bits, sandwiching_probability,
                            basis_probability, partner_layout, lookahead_period, lookahead_candidates, 0)
        self.ctx = ctypes.c_void_p()
        _check(self.L.fg_create(ctypes.byref(self.cfg), ctypes.byref(self.ctx)), "fg_create")
        self.walkers = walkers

    def close(self):
        if self.ctx:
            self.L.fg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_naive(self, n1, n2, n3):
        _check(self.L.fg_init_naive(self.ctx, n1, n2, n3), "fg_init_naive")

    def init_from_schemes(self, blob):
        b, p = _i32(blob)
        _check(self.L.fg_init_from_schemes(self.ctx, p, len(b)), "fg_init_from_schemes")

    def init_batched(self, formats, weights):
        f = np.asarray(formats, dtype=np.int32).reshape(-1, 3)
        w = np.asarray(weights, dtype=np.int32)
        f, pf = _i32(f.ravel())
        w, pw = _i32(w)
        _check(self.L.fg_init_batched(self.ctx, len(w), pf, pw), "fg_init_batched")

    def init_minimizer(self, blob):
        """Minimizer pool from a count-prefixed blob of schemes of one format."""
        b, p = _i32(blob)
        _check(self.L.fg_init_minimizer(self.ctx, p, len(b)), "fg_init_minimizer")

    def init_flips(self, blob):
        """Scheme pool of the additions reducer's --max-flips mode: every walker
        starts from the single integer-ring scheme `blob` (no count prefix)."""
        b = np.concatenate([[1], np.asarray(blob, dtype=np.int32)]).astype(np.int32)
        p = b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        _check(self.L.fg_init_flips(self.ctx, p, len(b)), "fg_init_flips")

    def flip_schemes(self, max_flips):
        """flipSchemesKernel: walkers >= 1 restart from walker 0's scheme and take
        up to max_flips flips; returns the number of schemes failing the Brent check."""
        return _check(self.L.fg_flip_schemes(self.ctx, int(max_flips)), "fg_flip_schemes")

    def resize(self, probability):
        """resizeKernel on every walker; returns the number of schemes failing the Brent check."""
        return _check(self.L.fg_resize(self.ctx, float(probability)), "fg_resize")

    def snapshot_best(self, walkers):
        """The current scheme of each listed walker becomes its best (updateRanks)."""
        w = np.ascontiguousarray(walkers, dtype=np.int32)
        _check(self.L.fg_snapshot_best(self.ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(w)),
               "fg_snapshot_best")

    def set_best_ranks(self, ranks):
        r = np.ascontiguousarray(ranks, dtype=np.int32)
        _check(self.L.fg_set_best_ranks(self.ctx, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "fg_set_best_ranks")

    def minimize(self):
        """One minimizer launch; returns the number of current schemes failing the GPU Brent check."""
        return _check(self.L.fg_minimize(self.ctx), "fg_minimize")

    def complexities(self, best=True):
        """(best complexities or None, current complexities) per walker."""
        W = self.walkers
        cur = np.zeros(W, dtype=np.int32)
        bst = np.zeros(W, dtype=np.int32) if best else None
        pb = bst.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if best else None
        _check(self.L.fg_complexities(self.ctx, pb, cur.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "fg_complexities")
        return bst, cur

    def broadcast_best(self, src, dst):
        d, p = _i32(dst)
        _check(self.L.fg_broadcast_best(self.ctx, int(src), p, len(d)), "fg_broadcast_best")

    # ---- learned proposals ----
    def log_enable(self, max_batches, every=1, horizon=10, window=0):
        """Trajectory log; with window > 0 (horizon 1) a flip is labelled by the
        first `window` iterations of its own launch."""
        _check(self.L.fg_log_enable(self.ctx, max_batches, every, horizon), "fg_log_enable")
        if window:
            _check(self.L.fg_log_window(self.ctx, window), "fg_log_window")

    def log_info(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(self.L.fg_log_info(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "fg_log_info")
        return {"rows": a.value, "positives": b.value, "batches": c.value}

    def log_read(self, begin, count, max_terms):
        words = np.zeros((count, 3, max_terms), dtype=np.uint32)
        m = np.zeros(count, dtype=np.int32)
        pairs = np.zeros((count, 3), dtype=np.int32)
        y = np.zeros(count, dtype=np.float32)
        _check(self.L.fg_log_read(self.ctx, begin, count, words.ctypes.data, m.ctypes.data, pairs.ctypes.data,
                                  y.ctypes.data), "fg_log_read")
        return words, m, pairs, y

    def set_forced(self, forced):
        f = np.ascontiguousarray(forced, dtype=np.int32)
        _check(self.L.fg_set_forced(self.ctx, f.ctypes.data), "fg_set_forced")

    def get_forced(self):
        f = np.zeros((self.walkers, 4), dtype=np.int32)
        _check(self.L.fg_get_forced(self.ctx, f.ctypes.data), "fg_get_forced")
        return f

    def walk_forced(self, iterations=0):
        return _check(self.L.fg_walk_forced(self.ctx, iterations), "fg_walk_forced")

    def walk_learned(self, scorer, iterations, candidates=8, every=2, round_=0):
        return _check(self.L.fg_walk_learned(self.ctx, scorer.s, iterations, candidates, every, round_),
                      "fg_walk_learned")

    def autotune(self, iterations):
        g = ctypes.c_int32()
        _check(self.L.fg_autotune(self.ctx, iterations, ctypes.byref(g)), "fg_autotune")
        return g.value

    LAYOUTS = {0: "z2 lists, 32-bit words", 1: "integer ring", 2: "z2 lists, 16-bit words", 3: "z2 class masks",
               4: "z2 class masks, one lane per walker"}

    def layout_info(self):
        """(layout, trials of the last autotune as [n][3] = layout, grid, flips/s / 1e6)."""
        v = ctypes.c_int32()
        buf = np.zeros((64, 3), dtype=np.int32)
        n = self.L.fg_layout_info(self.ctx, ctypes.byref(v), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 64)
        _check(n, "fg_layout_info")
        return int(v.value), buf[:n].copy()

    def summary(self):
        W = self.walkers
        arrs = [np.zeros(W, dtype=np.int32) for _ in range(4)] + [np.zeros(3 * W, dtype=np.int32)]
        ptrs = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) for a in arrs]
        _check(self.L.fg_summary(self.ctx, *ptrs), "fg_summary")
        return {"m": arrs[0], "best_m": arrs[1], "best_rank": arrs[2], "flips_count": arrs[3],
                "n": arrs[4].reshape(W, 3)}

    def walk(self, iterations=0):
        """One launch; returns the number of current schemes failing the GPU Brent check."""
        return _check(self.L.fg_walk(self.ctx, iterations), "fg_walk")

    def walk_trace(self, iterations=0, cap=4096):
        """fg_walk that also records every operator applied to each walker's
        scheme: returns (ops [walkers][cap][8] int32, counts [walkers], the
        number of schemes failing the GPU Brent check).  A count above cap
        means the walker's trace was cut."""
        W = self.walkers
        ops = np.zeros((W, max(cap, 1), 8), dtype=np.int32)
        counts = np.zeros(W, dtype=np.int32)
        bad = _check(self.L.fg_walk_trace(self.ctx, iterations, cap, ops.ctypes.data, counts.ctypes.data),
                     "fg_walk_trace")
        return ops[:, :cap], counts, bad

    def last_walk_ms(self):
        x = ctypes.c_float()
        _check(self.L.fg_last_walk_ms(self.ctx, ctypes.byref(x)), "fg_last_walk_ms")
        return x.value

    def counters(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(self.L.fg_counters(self.ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "fg_counters")
        return {"flips": a.value, "sum_m": b.value, "iterations": c.value}

    def read_walkers(self, begin=0, count=None):
        count = self.walkers - begin if count is None else count
        recs = np.zeros(count, dtype=RECORD_DTYPE)
        _check(self.L.fg_read_walkers(self.ctx, begin, count, recs.ctypes.data), "fg_read_walkers")
        return recs

    def write_walkers(self, recs, begin=0):
        recs = np.ascontiguousarray(recs, dtype=RECORD_DTYPE)
        _check(self.L.fg_write_walkers(self.ctx, begin, len(recs), recs.ctypes.data), "fg_write_walkers")

    def best(self, cap=1 << 16):
        out = np.zeros(cap, dtype=np.int32)
        rank, walker = ctypes.c_int32(), ctypes.c_int32()
        n = _check(self.L.fg_best(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                  ctypes.byref(rank), ctypes.byref(walker)), "fg_best")
        return out[:n], rank.value, walker.value

    def walker_scheme(self, walker, best=False, cap=1 << 16):
        out = np.zeros(cap, dtype=np.int32)
        n = _check(self.L.fg_walker_scheme(self.ctx, walker, int(best),
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap),
                   "fg_walker_scheme")
        return out[:n]


def verify_host(blobs, ring=RING_Z2):
    """The library's host Brent check (fg_verify_host) of single-scheme blobs; bool array."""
    L = load_library()
    flat = np.concatenate([np.array([len(blobs)], dtype=np.int32)] + [np.asarray(b, dtype=np.int32) for b in blobs])
    f, p = _i32(flat)
    ok = np.zeros(len(blobs), dtype=np.int32)
    _check(L.fg_verify_host(ring, p, len(f), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "fg_verify_host")
    return ok.astype(bool)


def complexity_host(blobs, ring=RING_Z2):
    """getComplexity of single-scheme blobs on the host (fg_complexity_host)."""
    L = load_library()
    flat = np.concatenate([np.array([len(blobs)], dtype=np.int32)] + [np.asarray(b, dtype=np.int32) for b in blobs])
    f, p = _i32(flat)
    out = np.zeros(len(blobs), dtype=np.int32)
    _check(L.fg_complexity_host(ring, p, len(f), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
           "fg_complexity_host")
    return out


def verify(blobs, ring=RING_Z2, noalloc=False):
    """GPU Brent check of a list of single-scheme blobs; returns a bool array.
    noalloc: through fg_verify_noalloc (staging in the library's fixed scratch)."""
    L = load_library()
    flat = np.concatenate([np.array([len(blobs)], dtype=np.int32)] + [np.asarray(b, dtype=np.int32) for b in blobs])
    f, p = _i32(flat)
    ok = np.zeros(len(blobs), dtype=np.int32)
    fn = L.fg_verify_noalloc if noalloc else L.fg_verify
    _check(fn(ring, p, len(f), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
           "fg_verify_noalloc" if noalloc else "fg_verify")
    return ok.astype(bool)


def _parse_part(buf):
    """Decode one serialised reducer component (fg_reducer_part layout)."""
    count, real, fresh, naive, mode, sbits, additions = (int(x) for x in buf[:7])
    p = 7
    vars_ = [(int(buf[p + 2 * v]), int(buf[p + 2 * v + 1])) for v in range(fresh)]
    p += 2 * fresh
    expr = []
    for _ in range(count):
        n = int(buf[p])
        expr.append([int(x) for x in buf[p + 1:p + 1 + n]])
        p += 1 + n
    scale = float(np.array([sbits], dtype=np.int32).view(np.float32)[0])
    return {"count": count, "real": real, "fresh": fresh, "naive": naive, "mode": mode, "scale": scale,
            "additions": additions, "vars": vars_, "expr": expr}


class Reducer:
    """Batched additions reducer on the current GPU (fg_reducer)."""

    def __init__(self, blob, count=32, schemes_count=1, max_flips=0, seed=1):
        self.L = load_library()
        b, p = _i32(blob)
        self.r = ctypes.c_void_p()
        _check(self.L.fg_reducer_create(p, len(b), count, schemes_count, max_flips, seed, ctypes.byref(self.r)),
               "fg_reducer_create")
        self.count = count

    def close(self):
        if self.r:
            self.L.fg_reducer_destroy(self.r)
            self.r = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(16, dtype=np.int32)
        _check(self.L.fg_reducer_info(self.r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "fg_reducer_info")
        keys = ["n1", "n2", "n3", "m", "count", "schemes_count", "max_flips", "top", "e_cap", "l_cap", "f_cap",
                "p_cap", "h_bits", "lds"]
        return {k: int(v) for k, v in zip(keys, out)}

    def run(self, iteration=1):
        _check(self.L.fg_reducer_run(self.r, iteration), "fg_reducer_run")

    def update(self):
        x = ctypes.c_int32()
        _check(self.L.fg_reducer_update(self.r, ctypes.byref(x)), "fg_reducer_update")
        return x.value

    def best(self):
        out = np.zeros(11, dtype=np.int32)
        _check(self.L.fg_reducer_best(self.r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "fg_reducer_best")
        return {"additions": out[0:3].tolist(), "fresh": out[3:6].tolist(), "naive": out[6:9].tolist(),
                "total": int(out[9]), "total_fresh": int(out[10])}

    def part(self, reducer, component, cap=1 << 20):
        out = np.zeros(cap, dtype=np.int32)
        n = self.L.fg_reducer_part(self.r, reducer, component, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap)
        _check(int(n), "fg_reducer_part")
        return _parse_part(out[:n])

    def best_part(self, component, cap=1 << 20):
        out = np.zeros(cap, dtype=np.int32)
        n = self.L.fg_reducer_best_part(self.r, component, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap)
        _check(int(n), "fg_reducer_best_part")
        return _parse_part(out[:n])

    def words(self):
        out = np.zeros(self.count * 4, dtype=np.uint64)
        _check(self.L.fg_reducer_words(self.r, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))), "fg_reducer_words")
        return out


class Scorer:
    """Learned flip scorer (fg_scorer): BF16 term-set network on the MFMA units."""

    def __init__(self, n1, n2, n3, max_terms, seed=1):
        self.L = load_library()
        self.s = ctypes.c_void_p()
        _check(self.L.fg_scorer_create(n1, n2, n3, max_terms, seed, ctypes.byref(self.s)), "fg_scorer_create")
        self.max_terms = max_terms
        self.nparams = int(self.L.fg_scorer_num_params(self.s))

    def close(self):
        if self.s:
            self.L.fg_scorer_destroy(self.s)
            self.s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self):
        out = np.zeros(16, dtype=np.int32)
        _check(self.L.fg_scorer_layout(self.s, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "fg_scorer_layout")
        return out

    def get_params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        _check(self.L.fg_scorer_get_params(self.s, out.ctypes.data), "fg_scorer_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.L.fg_scorer_set_params(self.s, p.ctypes.data), "fg_scorer_set_params")

    def eval(self, words, m, pairs, labels=None):
        """(scores, loss, grads); loss and grads only with labels."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        m = np.ascontiguousarray(m, dtype=np.int32)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        n = len(m)
        scores = np.zeros(n, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        grads = np.zeros(self.nparams, dtype=np.float32) if labels is not None else None
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.float32)
        _check(self.L.fg_scorer_eval(self.s, n, words.ctypes.data, m.ctypes.data, pairs.ctypes.data,
                                     None if lab is None else lab.ctypes.data, scores.ctypes.data,
                                     None if lab is None else loss.ctypes.data,
                                     None if lab is None else grads.ctypes.data), "fg_scorer_eval")
        return scores, (float(loss[0]) if lab is not None else None), grads

    def train(self, engine, steps, batch=256, lr=1e-3, seed=1):
        losses = np.zeros(steps, dtype=np.float32)
        _check(self.L.fg_scorer_train(self.s, engine.ctx, steps, batch, lr, seed, losses.ctypes.data),
               "fg_scorer_train")
        return losses

    def propose(self, engine, candidates=8, every=2, round_=0):
        _check(self.L.fg_scorer_propose(self.s, engine.ctx, candidates, every, round_), "fg_scorer_propose")

    def last_proposals(self, max_tuples, candidates):
        """(walkers [B], pairs [B][K][3], scores [B][K]) of the last fused proposal round."""
        cap = max_tuples * candidates
        wk = np.zeros(max_tuples, dtype=np.int32)
        pr = np.zeros((cap, 3), dtype=np.int32)
        sc = np.zeros(cap, dtype=np.float32)
        self.L.fg_scorer_last_proposals.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64]
        B = _check(self.L.fg_scorer_last_proposals(self.s, wk.ctypes.data, pr.ctypes.data, sc.ctypes.data, cap),
                   "fg_scorer_last_proposals")
        return wk[:B], pr[:B * candidates].reshape(B, candidates, 3), sc[:B * candidates].reshape(B, candidates)

    def last_propose_ms(self):
        """Duration (ms, HIP events) of the fused proposal kernel of the last round; 0 if it was not fused."""
        ms = ctypes.c_float()
        self.L.fg_scorer_last_propose_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        _check(self.L.fg_scorer_last_propose_ms(self.s, ctypes.byref(ms)), "fg_scorer_last_propose_ms")
        return float(ms.value)
