"""GPU additions reducer (fg_reducer_*) vs the oracle's restatement of
runReducersKernel (src/entities/scheme_additions_reducer.cu:485-540): every
reducer's mode, scale, fresh variables and reduced expressions must be
identical, over two passes with the best reducers replayed in the second,
and updateBest must pick what the reference's rules pick."""
import numpy as np
import pytest

import flipgraphgpu_amd as fg
import oracle_ffi as O
import reducer_ffi as RF
import schemes_np as S
from minimizer_util import walked_schemes

pytestmark = pytest.mark.gpu


def _inputs():
    n, U, V, W = S.strassen_squared()
    yield "strassen2", S.to_blob(n, U, V, W)
    yield "zt333", walked_schemes((3, 3, 3), 8, 4000, 3, O.RING_ZT)[0]
    yield "zt444", walked_schemes((4, 4, 4), 8, 3000, 3, O.RING_ZT)[0]


def _same(g, o):
    keys = ["count", "real", "fresh", "naive", "mode", "additions", "vars", "expr"]
    for k in keys:
        if g[k] != o[k]:
            return f"{k}: gpu={str(g[k])[:200]} oracle={str(o[k])[:200]}"
    if np.float32(g["scale"]) != np.float32(o["scale"]):
        return f"scale {g['scale']} vs {o['scale']}"
    return None


@pytest.mark.parametrize("name,blob", list(_inputs()), ids=lambda x: x if isinstance(x, str) else "")
def test_reducer_passes_bit_identical_to_oracle(name, blob):
    count, seed = 96, 77
    print(f"[reducer] {name} m={int(blob[3])} count={count}", flush=True)
    fixed = O.replay(blob, [[5]], ring=O.RING_ZT)  # the reader applies fixSigns
    with fg.Reducer(blob, count=count, seed=seed) as r:
        inf = r.info()
        best = None
        words = None
        for it in (1, 2):
            r.run(it)
            gw = r.words()
            ores, words = RF.iteration([fixed], count, seed, independent=True, best=best, words=words,
                                       e_cap=inf["e_cap"], l_cap=inf["l_cap"], f_cap=inf["f_cap"])
            assert np.array_equal(gw, words), "random stream positions differ"
            for q in range(count):
                for c in range(3):
                    msg = _same(r.part(q, c), ores[q][c])
                    assert msg is None, f"pass {it} reducer {q} component {c}: {msg}"
                    rows = RF.component_rows(fixed, c)
                    assert RF.expand(ores[q][c], rows.shape[1]) == RF.rows_as_forms(rows)
            improved = r.update()
            b = r.best()
            # updateBestIndependent: the best (additions, fresh) per component
            for c in range(3):
                top = min((ores[q][c]["additions"], ores[q][c]["fresh"]) for q in range(count))
                assert (b["additions"][c], b["fresh"][c]) <= top or it == 2
                bp = r.best_part(c)
                assert (bp["additions"], bp["fresh"]) == (b["additions"][c], b["fresh"][c])
            assert improved in (0, 1, 2) and (it == 2 or improved == 1)
            best = [(r.best_part(c)["fresh"], r.best_part(c)["vars"]) for c in range(3)]
        assert b["total"] < sum(b["naive"])


def test_reducer_rejects_invalid_scheme():
    n, U, V, W = S.strassen()
    blob = S.to_blob(n, U, V, W).copy()
    blob[5] = 1 - blob[5]
    with pytest.raises(fg.FgError):
        fg.Reducer(blob, count=4)
    with pytest.raises(fg.FgError):
        fg.Reducer(S.to_blob(n, U, V, W), count=4, schemes_count=0, max_flips=3)
