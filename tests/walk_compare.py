"""Helpers comparing GPU walker records with oracle records (bit-exact)."""
import numpy as np

SCALARS = ["m", "best_m", "best_rank", "flips_count", "aux_idx", "it", "aux", "total_flips", "sum_m"]


def first_mismatch(gpu, ora):
    """Return a description of the first differing walker, or None."""
    assert len(gpu) == len(ora)
    for w in range(len(gpu)):
        g, o = gpu[w], ora[w]
        if not np.array_equal(g["n"], o["n"]):
            return f"walker {w}: n {g['n']} vs {o['n']}"
        for f in SCALARS:
            if g[f] != o[f]:
                return f"walker {w}: {f} gpu={g[f]} oracle={o[f]}"
        m, bm = int(o["m"]), int(o["best_m"])
        for f, k in (("cur_v", m), ("cur_s", m), ("best_v", bm), ("best_s", bm)):
            if not np.array_equal(g[f][:, :k], o[f][:, :k]):
                return f"walker {w}: {f} differs"
    return None
