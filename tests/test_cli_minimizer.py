"""The complexity_minimizer CLI keeps the reference's flags
(src/main_complexity_minimizer.cu:8-31), input format
(src/entities/complexity_minimizer.cu:22-37), report table and stop rule
(:39-65, :103-128); GPU runs check the saved schemes' validity, complexity
and names (:136-149)."""
import glob
import json
import os
import re
import subprocess

import pytest

import oracle_ffi as O
import schemes_np as S
from minimizer_util import walked_schemes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "flipgraphgpu_amd", "bin", "complexity_minimizer")


def run(*args, timeout=60):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)


def write_input(path, schemes):
    n1, n2, n3, m = (int(x) for x in schemes[0][:4])
    lines = [f"{n1} {n2} {n3} {m} {len(schemes)}"]
    for b in schemes:
        lines.append(" ".join(str(int(x)) for x in b[4:]))
    path.write_text("\n".join(lines) + "\n")


def test_help_lists_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ["--schemes 1024", "--max-iterations 10000", "--path schemes", "--input-path PATH",
                 "--block-size 32", "--target-complexity 0", "--max-no-improvements 3", "--seed 0"]:
        assert flag in r.stdout, flag


@pytest.mark.parametrize("args,msg", [
    ([], 'no value for argument "--input-path"'),
    (["--input-path"], "Invalid number of arguments"),
    (["--input-path", "x", "--schemes", "-4"], 'value for arg "--schemes" is not natural'),
    (["--input-path", "/nonexistent/file.txt"], 'Unable to open file "/nonexistent/file.txt"'),
])
def test_parse_errors(args, msg):
    r = run(*args)
    assert msg in r.stdout + r.stderr


def test_truncated_input_is_rejected(tmp_path):
    # the header promises two schemes, the file holds one
    blob = O.naive_blob(2, 2, 2)
    p = tmp_path / "in.txt"
    p.write_text("2 2 2 8 2\n" + " ".join(str(int(x)) for x in blob[4:]) + "\n")
    r = run("--input-path", str(p))
    assert r.returncode != 0
    assert "Read invalid scheme" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("ring", ["Z2", "ZT"])
def test_minimizer_improves_and_saves(tmp_path, ring):
    print(f"[cli] complexity_minimizer {ring}", flush=True)
    rid = O.RING_ZT if ring == "ZT" else O.RING_Z2
    schemes = walked_schemes((3, 3, 3), 16, 2000, 11, rid)
    cx0 = min(O.complexity(b, rid) for b in schemes)
    m = int(schemes[0][3])
    inp = tmp_path / "in.txt"
    write_input(inp, schemes)
    out = tmp_path / "schemes"
    r = run("--input-path", str(inp), "--schemes", "512", "--max-iterations", "2000", "--max-no-improvements", "2",
            "--seed", "5", "--path", str(out), "--ring", ring, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Start reading {len(schemes)} schemes (3x3x3: {m})" in r.stdout
    assert "|  elapsed  | iteration | run id |  best  |  curr  |" in r.stdout
    assert "No improvements for 2 iterations" in r.stdout and r.stdout.rstrip().endswith("Success!")
    gains = [(int(a), int(b)) for a, b in re.findall(r"Best complexity improved from (\d+) to (\d+)!", r.stdout)]
    assert gains and gains[0][0] == cx0
    assert all(b < a for a, b in gains) and all(gains[q][1] == gains[q + 1][0] for q in range(len(gains) - 1))
    files = glob.glob(str(out / "*.json"))
    assert len(files) == len(gains)
    for f in files:
        name = os.path.basename(f)
        mm = re.fullmatch(r"3x3x3_m(\d+)_c(\d+)_iteration\d+_run\d+_3x3x3_" + ring + r"\.json", name)
        assert mm, name
        d = json.load(open(f))
        blob = S.to_blob((3, 3, 3), d["u"], d["v"], d["w"])
        assert O.validate(blob, rid) == 1
        assert d["complexity"] == O.complexity(blob, rid) == int(mm.group(2))
    best = min(b for _, b in gains)
    assert any(f"_c{best}_" in os.path.basename(f) for f in files)
