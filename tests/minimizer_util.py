"""Record-level helpers for the complexity-minimizer tests: the start state
of ComplexityMinimizer::initialize (src/entities/complexity_minimizer.cu:67-77,
172-188) and its re-broadcast (:130-133), applied to walker records."""
import numpy as np

import oracle_ffi as O


def record_complexity(rec, ring=O.RING_Z2):
    """getComplexity (scheme_z2.cu:127-139): coefficient count - 2m - n3*n1."""
    m = int(rec["m"])
    n1, _, n3 = (int(x) for x in rec["n"])
    vals = rec["cur_v"][:, :m].astype(np.uint64)
    if ring == O.RING_ZT:
        vals = vals & np.uint64(0xFFFF)
    pop = sum(bin(int(x)).count("1") for x in vals.ravel())
    return pop - 2 * m - n3 * n1


def minimizer_start(recs, ring=O.RING_Z2):
    """best := current, best complexity := the lowest current complexity."""
    out = np.array(recs, dtype=O.RECORD_DTYPE, copy=True)
    cx = min(record_complexity(r, ring) for r in out)
    out["best_v"] = out["cur_v"]
    out["best_s"] = out["cur_s"]
    out["best_m"] = out["m"]
    out["best_rank"] = cx
    out["flips_count"] = 0
    return out, cx


def broadcast(recs, src, dst):
    """schemesBest[src].copyTo(schemes[d]) for every d in dst."""
    out = np.array(recs, dtype=O.RECORD_DTYPE, copy=True)
    for d in dst:
        out[d]["cur_v"] = recs[src]["best_v"]
        out[d]["cur_s"] = recs[src]["best_s"]
        out[d]["m"] = recs[src]["best_m"]
    return out


def rebroadcast_targets(indices, iteration, period):
    """Walkers receiving the best scheme after a report (:130-133)."""
    return [indices[i] for i in range(len(indices)) if i % (iteration % period + 1) == 0]


def record_blob(rec, ring=O.RING_Z2, best=False):
    """Single-scheme blob (n1 n2 n3 m U V W) of a record's current or best scheme."""
    n1, n2, n3 = (int(x) for x in rec["n"])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    m = int(rec["best_m"] if best else rec["m"])
    vals = rec["best_v" if best else "cur_v"]
    sgns = rec["best_s" if best else "cur_s"]
    out = [n1, n2, n3, m]
    for c in range(3):
        for t in range(m):
            v, s = int(vals[c][t]), int(sgns[c][t])
            for j in range(nn[c]):
                x = (v >> j) & 1
                out.append(-x if ring == O.RING_ZT and (s >> j) & 1 else x)
    return np.array(out, dtype=np.int32)


def walked_schemes(n, count, iterations, seed, ring=O.RING_Z2, same_rank=True):
    """Schemes a few thousand walk iterations away from the naive scheme, as
    blobs; with same_rank only those of the most common rank."""
    recs = O.naive_records(*n, count, ring=ring)
    recs = O.walk_records(recs, O.make_params(iterations=iterations, plus_iterations=400), seed, ring=ring,
                          nthreads=8)
    blobs = [record_blob(r, ring) for r in recs]
    if same_rank:
        ranks = [int(b[3]) for b in blobs]
        common = max(set(ranks), key=ranks.count)
        blobs = [b for b in blobs if int(b[3]) == common]
    return blobs


def records_from_blob(blob, count, ring=O.RING_ZT):
    """`count` walker records whose current and best scheme is `blob` (coefficients
    in {-1, 0, 1}), with fresh random streams: the state fg_init_flips sets up."""
    n1, n2, n3, m = (int(x) for x in blob[:4])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    recs = np.zeros(count, dtype=O.RECORD_DTYPE)
    recs["n"] = (n1, n2, n3)
    recs["m"] = m
    recs["best_m"] = m
    recs["best_rank"] = m
    recs["aux_idx"] = 4
    p = 4
    for c in range(3):
        for t in range(m):
            row = [int(x) for x in blob[p:p + nn[c]]]
            p += nn[c]
            v = sum(1 << j for j, x in enumerate(row) if x)
            s = sum(1 << j for j, x in enumerate(row) if x < 0) if ring == O.RING_ZT else 0
            recs["cur_v"][:, c, t] = v
            recs["cur_s"][:, c, t] = s
            recs["best_v"][:, c, t] = v
            recs["best_s"][:, c, t] = s
    return recs


def flip_launches(blob, count, max_flips, seed, launches, run):
    """Oracle-side driver of the additions reducer's scheme flips: before every
    launch walkers 1.. restart from walker 0's scheme (its best, which it never
    leaves); `run(recs, params)` performs one flips launch.  Returns the
    records after each launch."""
    recs = records_from_blob(blob, count)
    params = O.make_params(iterations=max_flips, mode=3, zt_visit_cap=16)
    out = []
    for _ in range(launches):
        recs = broadcast(recs, 0, range(1, count))
        recs = run(recs, params)
        out.append(recs)
    return out
