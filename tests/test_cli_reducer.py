"""The additions_reducer CLI keeps the reference's flags
(src/main_additions_reducer.cu:9-20), input format, stop rule, report and
JSON output (src/entities/scheme_additions_reducer.cu:103-130, 365-401,
src/entities/additions_reducer.cuh:296-347); GPU runs check that the saved
expressions, with the fresh variables substituted, are the scheme's."""
import glob
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
import reducer_ffi as RF
import schemes_np as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "flipgraphgpu_amd", "bin", "additions_reducer")


def run(*args, timeout=60):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)


def test_help_lists_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ["-i PATH", "-o schemes/additions_reduced", "--count 32", "--schemes-count 1", "--block-size 32",
                 "--start-additions 0", "--max-flips 0", "--max-no-improvements 3", "--seed 0"]:
        assert flag in r.stdout, flag


@pytest.mark.parametrize("args,msg", [
    ([], 'no value for argument "-i"'),
    (["-i", "x", "--count", "1.5"], 'value for arg "--count" is not natural'),
    (["-i", "/nonexistent/s.txt"], 'Unable to open file "/nonexistent/s.txt"'),
])
def test_parse_errors(args, msg):
    r = run(*args)
    assert msg in r.stdout + r.stderr


def _json_forms(part, fresh, n_real):
    """Expand a saved component: fresh variables substituted into each expression."""
    res = {"real": n_real, "vars": [], "expr": []}
    for a, b in fresh:
        res["vars"].append(tuple((x["index"] + 1) * x["value"] for x in (a, b)))
    for e in part:
        res["expr"].append([(x["index"] + 1) * x["value"] for x in e])
    return RF.expand(res, n_real)


@pytest.mark.gpu
def test_reducer_saves_valid_reductions(tmp_path):
    print("[cli] additions_reducer strassen2", flush=True)
    n, U, V, W = S.strassen_squared()
    blob = S.to_blob(n, U, V, W)
    inp = tmp_path / "s.txt"
    inp.write_text(" ".join(str(int(x)) for x in blob[:4]) + "\n" + " ".join(str(int(x)) for x in blob[4:]) + "\n")
    out = tmp_path / "red"
    r = run("-i", str(inp), "-o", str(out), "--count", "128", "--seed", "3", "--max-no-improvements", "2",
            timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "- naive additions (U / V / W / total): 95 / 95 / 128 / 318" in r.stdout
    assert "| mode | naive | reduced | fresh |" in r.stdout
    assert "No improvements for 2 iterations" in r.stdout
    gains = [int(b) for b in re.findall(r"Reduced scheme improved from \d+ to (\d+) additions", r.stdout)]
    assert gains and min(gains) < 318
    files = sorted(glob.glob(str(out / "*.json")))
    assert files
    fixed = O.replay(blob, [[5]], ring=O.RING_ZT)
    for f in files:
        mm = re.fullmatch(r"4x4x4_m49_cr(\d+)_fv(\d+)_cn318_ZT_reduced\.json", os.path.basename(f))
        assert mm, f
        d = json.load(open(f))
        assert d["n"] == [4, 4, 4] and d["m"] == 49 and d["z2"] is False
        assert d["complexity"] == {"naive": 318, "reduced": int(mm.group(1))}
        total = fresh = 0
        for c, name in enumerate("uvw"):
            rows = RF.component_rows(fixed, c)
            forms = _json_forms(d[name], d[name + "_fresh"], rows.shape[1])
            assert forms == RF.rows_as_forms(rows), f"{f}: {name}"
            fresh += len(d[name + "_fresh"])
            total += len(d[name + "_fresh"]) + sum(len(e) - 1 for e in d[name])
        assert total == int(mm.group(1)) and fresh == int(mm.group(2))


@pytest.mark.gpu
def test_reducer_with_scheme_flips_saves_valid_schemes(tmp_path):
    # --max-flips > 0: the reducers read flipped copies of the input
    # (flipSchemesKernel, scheme_additions_reducer.cu:448-460) and the best is
    # taken over whole reducers (updateBestTogether, :253-285), so each saved
    # file is the reduced form of one valid, possibly flipped, scheme
    from minimizer_util import walked_schemes
    print("[cli] additions_reducer --max-flips", flush=True)
    blob = walked_schemes((3, 3, 3), 8, 4000, 3, O.RING_ZT)[0]
    m = int(blob[3])
    inp = tmp_path / "s.txt"
    inp.write_text(" ".join(str(int(x)) for x in blob[:4]) + "\n" + " ".join(str(int(x)) for x in blob[4:]) + "\n")
    out = tmp_path / "red"
    r = run("-i", str(inp), "-o", str(out), "--count", "64", "--schemes-count", "8", "--max-flips", "5",
            "--seed", "4", "--max-no-improvements", "2", timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "- schemes count: 8" in r.stdout and "- max flips: 5" in r.stdout
    assert "No improvements for 2 iterations" in r.stdout
    files = sorted(glob.glob(str(out / "*.json")))
    assert files
    for f in files:
        d = json.load(open(f))
        mm = len(d["u"])  # rank of the scheme the best reducer read
        mats = [_json_forms(d[name], d[name + "_fresh"], (9, 9, mm)[c]) for c, name in enumerate("uvw")]
        assert len(mats[1]) == mm and mm <= m
        U = np.array([[fm.get(k + 1, 0) for k in range(9)] for fm in mats[0]])
        V = np.array([[fm.get(k + 1, 0) for k in range(9)] for fm in mats[1]])
        W = np.array([[fm.get(t + 1, 0) for t in range(mm)] for fm in mats[2]]).T
        assert S.brent_ok((3, 3, 3), U, V, W, ring="ZT"), f
