"""json_to_input turns saved scheme JSON (the layout of
src/schemes/scheme_z2.cu:1327-1364 and scheme_integer.cu:1247-1274) back into
the input text of the three tools (src/entities/flip_graph.cu:201-212,
complexity_minimizer.cu:22-37, scheme_additions_reducer.cu:29-48)."""
import glob
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
import schemes_np as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "flipgraphgpu_amd", "bin", "json_to_input")


def save_json(path, blob, z2):
    n1, n2, n3, m = (int(x) for x in blob[:4])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    p, mats = 4, []
    for c in range(3):
        mats.append([[int(x) for x in blob[p + t * nn[c]:p + (t + 1) * nn[c]]] for t in range(m)])
        p += m * nn[c]
    d = {"n": [n1, n2, n3], "m": m, "z2": z2, "complexity": O.complexity(blob, 0 if z2 else 1),
         "u": mats[0], "v": mats[1], "w": mats[2]}
    with open(path, "w") as f:
        f.write(json.dumps(d, indent=4))


def ints(path):
    return [int(x) for x in open(path).read().split()]


@pytest.mark.parametrize("fmt", ["flip_graph", "complexity_minimizer", "additions_reducer"])
def test_round_trip(tmp_path, fmt):
    blobs = [S.to_blob(*S.strassen()), O.naive_blob(2, 2, 2)] if fmt == "flip_graph" else [S.to_blob(*S.strassen())]
    if fmt == "complexity_minimizer":
        n, U, V, W = S.strassen()
        blobs.append(S.to_blob(n, U[::-1], V[::-1], W[::-1]))  # same rank, terms reversed
    paths = []
    for q, b in enumerate(blobs):
        p = tmp_path / f"s{q}.json"
        save_json(p, b, z2=False)
        paths.append(str(p))
    out = tmp_path / "in.txt"
    r = subprocess.run([BIN, "-i", ",".join(paths), "-o", str(out), "--format", fmt], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    got = ints(out)
    if fmt == "flip_graph":
        want = [len(blobs)] + [int(x) for b in blobs for x in b]
    elif fmt == "complexity_minimizer":
        want = [int(x) for x in blobs[0][:4]] + [len(blobs)] + [int(x) for b in blobs for x in b[4:]]
    else:
        want = [int(x) for x in blobs[0]]
    assert got == want


def test_rejects_reducer_output_and_mixed_ranks(tmp_path):
    p = tmp_path / "r.json"
    p.write_text('{"n": [2, 2, 2], "m": 7, "z2": false, "complexity": {"naive": 18, "reduced": 15}, '
                 '"u_fresh": [], "u": [[{"index": 0, "value": 1}]]}')
    r = subprocess.run([BIN, "-i", str(p), "-o", str(tmp_path / "x.txt")], capture_output=True, text=True)
    assert r.returncode != 0 and "expressions" in r.stdout
    a, b = tmp_path / "a.json", tmp_path / "b.json"
    save_json(a, S.to_blob(*S.strassen()), False)
    save_json(b, O.naive_blob(2, 2, 2), False)
    r = subprocess.run([BIN, "-i", f"{a},{b}", "-o", str(tmp_path / "y.txt"), "--format", "complexity_minimizer"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "one format and rank" in r.stdout


def _tool(name, *args, timeout=120):
    return subprocess.run([os.path.join(ROOT, "flipgraphgpu_amd", "bin", name), *args], capture_output=True,
                          text=True, timeout=timeout)


@pytest.mark.gpu
def test_saved_schemes_resume_all_three_tools(tmp_path):
    # flip_graph finds and saves rank-7 2x2x2 schemes (Z2); converted back,
    # they seed flip_graph and complexity_minimizer again.
    print("[convert] flip_graph 2x2x2 search", flush=True)
    r = _tool("flip_graph", "-n1", "2", "-n2", "2", "-n3", "2", "--schemes", "256", "--max-iterations", "3000",
              "--resize-probability", "0", "--max-launches", "3", "--target-rank", "7", "--seed", "1",
              "--path", str(tmp_path / "fg"))
    assert r.returncode == 0, r.stdout + r.stderr
    saved = sorted(glob.glob(str(tmp_path / "fg" / "*_m7_*.json")))
    assert saved
    txt = tmp_path / "fg.txt"
    r = _tool("json_to_input", "-i", ",".join(saved), "-o", str(txt))
    assert r.returncode == 0, r.stdout
    print("[convert] flip_graph resume", flush=True)
    r = _tool("flip_graph", "-n1", "2", "-n2", "2", "-n3", "2", "--schemes", "64", "--max-iterations", "200",
              "--resize-probability", "0", "--max-launches", "1", "--input-path", str(txt), "--seed", "2",
              "--path", str(tmp_path / "fg2"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Start reading {len(saved)} schemes" in r.stdout
    assert re.search(r"\|\s+\d+ \|  2x2x2 \|  2x2x2 \|     7 \|    7 \|", r.stdout), r.stdout[-2000:]
    cm = tmp_path / "cm.txt"
    r = _tool("json_to_input", "-i", ",".join(saved), "-o", str(cm), "--format", "complexity_minimizer")
    assert r.returncode == 0, r.stdout
    print("[convert] complexity_minimizer Z2 resume", flush=True)
    r = _tool("complexity_minimizer", "--input-path", str(cm), "--schemes", "64", "--max-iterations", "500",
              "--max-no-improvements", "1", "--seed", "4", "--path", str(tmp_path / "cm"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Start reading {len(saved)} schemes (2x2x2: 7)" in r.stdout

    # A ZT scheme saved by complexity_minimizer feeds additions_reducer.
    n, U, V, W = S.strassen_squared()
    blob = S.to_blob(n, U, V, W)
    (tmp_path / "st.txt").write_text("4 4 4 49 1\n" + " ".join(str(int(x)) for x in blob[4:]) + "\n")
    print("[convert] complexity_minimizer ZT strassen^2", flush=True)
    r = _tool("complexity_minimizer", "--input-path", str(tmp_path / "st.txt"), "--schemes", "64",
              "--max-iterations", "300", "--max-no-improvements", "1", "--seed", "6", "--ring", "ZT",
              "--path", str(tmp_path / "zt"))
    assert r.returncode == 0, r.stdout + r.stderr
    zt = sorted(glob.glob(str(tmp_path / "zt" / "*_ZT.json")))
    src = zt[0] if zt else None
    if src is None:  # no improvement saved: convert the input scheme itself
        src = str(tmp_path / "st.json")
        save_json(src, blob, z2=False)
    cx = json.load(open(src))["complexity"]
    red_in = tmp_path / "red.txt"
    r = _tool("json_to_input", "-i", src, "-o", str(red_in), "--format", "additions_reducer")
    assert r.returncode == 0, r.stdout
    print("[convert] additions_reducer", flush=True)
    r = _tool("additions_reducer", "-i", str(red_in), "-o", str(tmp_path / "red"), "--count", "32", "--seed", "3",
              "--max-no-improvements", "1")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "- multiplications (rank): 49" in r.stdout
    assert re.search(r"- naive additions \(U / V / W / total\): \d+ / \d+ / \d+ / %d\n" % cx, r.stdout), r.stdout
