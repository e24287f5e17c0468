"""The additions reducer's scheme flips (--max-flips > 0): flipSchemesKernel
(src/entities/scheme_additions_reducer.cu:448-460) restarts every scheme but
scheme 0 from scheme 0 and applies a random number of flips with reductions
before each pass, and the reducers then read scheme r % schemes-count with
one shared replay decision (:519-528) and updateBestTogether (:253-285).

The flip pool runs the walk kernel in its flips mode; the host emulation of
that kernel code, the GPU and the oracle must give bit-identical schemes and
random-stream positions, and the GPU reducer must match the oracle reducer
on the flipped schemes."""
import numpy as np
import pytest

import emu_ffi as E
import oracle_ffi as O
import reducer_ffi as RF
import schemes_np as S
from minimizer_util import flip_launches, record_blob, walked_schemes

FIELDS = ["m", "it", "aux", "aux_idx", "total_flips", "sum_m"]


def _inputs():
    n, U, V, W = S.strassen_squared()
    yield "strassen2", O.replay(S.to_blob(n, U, V, W), [[5]], ring=O.RING_ZT)
    yield "zt333", O.replay(walked_schemes((3, 3, 3), 8, 4000, 3, O.RING_ZT)[0], [[5]], ring=O.RING_ZT)


def _same_records(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"
    for w in range(len(a)):
        m = int(a[w]["m"])
        assert np.array_equal(a[w]["cur_v"][:, :m], b[w]["cur_v"][:, :m]), f"{what}: walker {w} factors"
        assert np.array_equal(a[w]["cur_s"][:, :m], b[w]["cur_s"][:, :m]), f"{what}: walker {w} signs"


def _oracle_run(seed):
    return lambda recs, p: O.walk_records(recs, p, seed, ring=O.RING_ZT)


@pytest.mark.parametrize("name,blob", list(_inputs()), ids=lambda x: x if isinstance(x, str) else "")
def test_emulated_flips_match_oracle(name, blob):
    count, max_flips, seed = 40, 6, 11
    emu = flip_launches(blob, count, max_flips, seed, 3, lambda r, p: E.walk(r, p, seed, ring=O.RING_ZT)[0])
    ora = flip_launches(blob, count, max_flips, seed, 3, _oracle_run(seed))
    for q, (e, o) in enumerate(zip(emu, ora)):
        _same_records(e, o, f"{name} launch {q}")
        for w in range(count):
            assert O.record_valid(o, w, ring=O.RING_ZT) == 1
    # walker 0 keeps the input; the others move unless no two factors of a
    # component are equal (Strassen's square has none, so nothing flips)
    assert np.array_equal(record_blob(ora[-1][0], O.RING_ZT), blob)
    flips = int(ora[-1]["total_flips"][1:].sum())
    assert (flips > 0) == (name != "strassen2")
    # each launch restarts from walker 0: flips never accumulate beyond max_flips per launch
    per_launch = np.diff(np.stack([o["total_flips"] for o in ora]).astype(np.int64), axis=0)
    assert per_launch.max() <= max_flips


@pytest.mark.gpu
@pytest.mark.parametrize("name,blob", list(_inputs()), ids=lambda x: x if isinstance(x, str) else "")
def test_flip_pool_bit_identical_to_oracle(name, blob):
    import flipgraphgpu_amd as fg
    count, max_flips, seed = 200, 8, 5
    print(f"[flips] {name} m={int(blob[3])} walkers={count}", flush=True)
    ora = flip_launches(blob, count, max_flips, seed, 3, _oracle_run(seed))
    e = fg.Engine(count, seed, ring=fg.RING_ZT, max_iterations=max_flips)
    try:
        e.init_flips(blob)
        for q in range(3):
            assert e.flip_schemes(max_flips) == 0
            _same_records(e.read_walkers(), ora[q], f"{name} launch {q}")
        with pytest.raises(fg.FgError):
            e.walk(10)
    finally:
        e.close()


@pytest.mark.gpu
def test_reducer_flips_mode_bit_identical_to_oracle():
    import flipgraphgpu_amd as fg
    blob = walked_schemes((3, 3, 3), 8, 4000, 3, O.RING_ZT)[0]
    fixed = O.replay(blob, [[5]], ring=O.RING_ZT)
    count, schemes, max_flips, seed = 48, 6, 4, 9
    print(f"[flips] reducer zt333 m={int(blob[3])} count={count} schemes={schemes}", flush=True)
    ora = flip_launches(fixed, schemes, max_flips, seed, 2, _oracle_run(seed))
    with fg.Reducer(blob, count=count, schemes_count=schemes, max_flips=max_flips, seed=seed) as r:
        inf = r.info()
        assert inf["schemes_count"] == schemes and inf["max_flips"] == max_flips
        best, words = None, None
        for it in (1, 2, 3):
            r.run(it)
            blobs = [fixed] if it == 1 else [record_blob(x, O.RING_ZT) for x in ora[it - 2]]
            ores, words = RF.iteration(blobs, count, seed, independent=False, best=best, words=words,
                                       e_cap=inf["e_cap"], l_cap=inf["l_cap"], f_cap=inf["f_cap"])
            assert np.array_equal(r.words().reshape(-1, 4)[:, :3], words.reshape(-1, 4)[:, :3])
            for q in range(count):
                for c in range(3):
                    g, o = r.part(q, c), ores[q][c]
                    for k in ["count", "real", "fresh", "naive", "mode", "additions", "vars", "expr"]:
                        assert g[k] == o[k], f"pass {it} reducer {q} component {c}: {k}"
            r.update()
            b = r.best()
            # updateBestTogether: the best reducer by total additions, then fresh variables
            tot = [(sum(ores[q][c]["additions"] for c in range(3)), sum(ores[q][c]["fresh"] for c in range(3)))
                   for q in range(count)]
            assert (b["total"], b["total_fresh"]) <= min(tot)
            best = [(r.best_part(c)["fresh"], r.best_part(c)["vars"]) for c in range(3)]
        assert b["total"] < sum(b["naive"])
