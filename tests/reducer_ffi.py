"""ctypes access to the oracle's additions reducer (oracle/fg_reducer.c) and
helpers shared by the reducer tests.  Test infrastructure only."""
import ctypes

import numpy as np

import oracle_ffi as O

MODE_NAMES = {0: "g", 1: "ga", 2: "gr", 3: "gi", 4: "wr", 5: "mix", 6: "rnd"}
MAX_FRESH = (250, 250, 500)
_done = False


def lib():
    global _done
    L = O.lib()
    if not _done:
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.fgo_reducer_single.argtypes = [i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, i32p, i32p, i32p, i32p]
        L.fgo_reducer_iteration.argtypes = [i32p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                            i32p, i32p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p]
        _done = True
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def unpack(info, size, expr, vars_):
    """One (reducer, component) result as a dict of python values."""
    n = int(info[0])
    fresh = int(info[2])
    return {
        "count": n, "real": int(info[1]), "fresh": fresh, "naive": int(info[3]), "mode": int(info[4]),
        "scale": float(np.array([info[5]], dtype=np.int32).view(np.float32)[0]), "additions": int(info[6]),
        "words": int(info[7]),
        "expr": [[int(x) for x in expr[q, :size[q]]] for q in range(n)],
        "vars": [(int(vars_[v, 0]), int(vars_[v, 1])) for v in range(fresh)],
    }


def single(matrix, mode, seed=1, r=0, c=0, max_fresh=250):
    """Run one reducer over the rows of a {-1,0,1} matrix (one expression per row)."""
    m = np.ascontiguousarray(matrix, dtype=np.int32)
    ne, nv = m.shape
    e_cap, l_cap, f_cap = ne, max(nv, 1), max_fresh
    info = np.zeros(8, dtype=np.int32)
    size = np.zeros(e_cap, dtype=np.int32)
    expr = np.zeros((e_cap, l_cap), dtype=np.int32)
    vars_ = np.zeros((f_cap, 2), dtype=np.int32)
    rc = lib().fgo_reducer_single(_p(m), ne, nv, mode, seed, r, c, max_fresh, e_cap, l_cap, f_cap, _p(info),
                                  _p(size), _p(expr), _p(vars_))
    assert rc == 0
    return unpack(info, size, expr, vars_)


def caps(blob):
    """(e_cap, l_cap) covering the U, V and W expressions of a single-scheme blob."""
    n1, n2, n3, m = (int(x) for x in blob[:4])
    return max(m, n3 * n1), max(n1 * n2, n2 * n3, m)


def iteration(blobs, count, seed, independent=True, best=None, words=None, e_cap=None, l_cap=None, f_cap=500):
    """One runReducersKernel pass of `count` reducers over the given schemes.
    best: [(fresh, [(i, j), ...])] * 3 or None.  Returns (results[r][c], words)."""
    flat = np.concatenate([np.array([len(blobs)], dtype=np.int32)] + [np.asarray(b, dtype=np.int32) for b in blobs])
    ec, lc = caps(blobs[0])
    e_cap = e_cap or ec
    l_cap = l_cap or lc
    bf = np.zeros(3, dtype=np.int32)
    bv = np.zeros((3, f_cap, 2), dtype=np.int32)
    if best is not None:
        for c in range(3):
            bf[c] = best[c][0]
            for v, (i, j) in enumerate(best[c][1]):
                bv[c, v] = (i, j)
    w = np.zeros(count * 4, dtype=np.uint64) if words is None else np.array(words, dtype=np.uint64)
    info = np.zeros((count * 3, 8), dtype=np.int32)
    size = np.zeros((count * 3, e_cap), dtype=np.int32)
    expr = np.zeros((count * 3, e_cap, l_cap), dtype=np.int32)
    vars_ = np.zeros((count * 3, f_cap, 2), dtype=np.int32)
    rc = lib().fgo_reducer_iteration(_p(flat), len(flat), count, seed, int(independent), _p(bf), _p(bv),
                                     w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), e_cap, l_cap, f_cap,
                                     _p(info), _p(size), _p(expr), _p(vars_))
    assert rc == 0
    res = [[unpack(info[r * 3 + c], size[r * 3 + c], expr[r * 3 + c], vars_[r * 3 + c]) for c in range(3)]
           for r in range(count)]
    return res, w


def expand(res, n_real):
    """Linear forms (dicts var -> coefficient over the real variables 1..n_real)
    of every expression, with fresh variables substituted."""
    defs = {}
    for v, (i, j) in enumerate(res["vars"]):
        idx = res["real"] + v + 1
        form = {}
        for x in (i, j):
            sub = defs[abs(x)] if abs(x) > res["real"] else {abs(x): 1}
            s = 1 if x > 0 else -1
            for k, val in sub.items():
                form[k] = form.get(k, 0) + s * val
        defs[idx] = {k: val for k, val in form.items() if val}
    out = []
    for e in res["expr"]:
        form = {}
        for x in e:
            sub = defs[abs(x)] if abs(x) > res["real"] else {abs(x): 1}
            s = 1 if x > 0 else -1
            for k, val in sub.items():
                form[k] = form.get(k, 0) + s * val
        out.append({k: val for k, val in form.items() if val})
    return out


def component_rows(blob, c):
    """The {-1,0,1} rows a component's reducer starts from (U rows, V rows, W columns)."""
    n1, n2, n3, m = (int(x) for x in blob[:4])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    U = np.asarray(blob[4:4 + m * nn[0]]).reshape(m, nn[0])
    V = np.asarray(blob[4 + m * nn[0]:4 + m * (nn[0] + nn[1])]).reshape(m, nn[1])
    W = np.asarray(blob[4 + m * (nn[0] + nn[1]):4 + m * sum(nn)]).reshape(m, nn[2])
    return (U, V, W.T)[c]


def rows_as_forms(rows):
    return [{k + 1: int(v) for k, v in enumerate(row) if v} for row in rows]
