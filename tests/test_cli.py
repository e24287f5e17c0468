"""The flip_graph CLI keeps the reference's flag rules
(src/entities/arg_parser.cu:16-146, src/main_flip_graph.cu:10-60) and output
format; GPU runs check that saved schemes are valid and named as the
reference names them (src/entities/flip_graph.cu:386-399)."""
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
BIN = os.path.join(ROOT, "flipgraphgpu_amd", "bin", "flip_graph")


def run(*args, timeout=60):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)


def test_help_lists_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ["-n1", "-n2", "-n3", "--schemes 1024", "--max-iterations 10000", "--path schemes",
                 "--input-path null", "--block-size 32", "--expand-probability 0.01", "--reduce-probability 0.0",
                 "--sandwiching-probability 0.0", "--basis-probability 0.0", "--resize-probability 0.2",
                 "--plus-iterations 8000", "--log-period 0", "--seed 0"]:
        assert flag in r.stdout, flag


@pytest.mark.parametrize("args,msg", [
    (["-n1"], "Invalid number of arguments"),
    (["-n1", "2", "--bogus", "1"], 'unknown argument "--bogus"'),
    (["-n1", "2x", "-n2", "2", "-n3", "2"], 'value for arg "-n1" is not natural'),
    (["-n1", "2", "-n2", "2", "-n3", "2", "--expand-probability", "0.1.2"], "is not real"),
    (["-n1", "2", "-n2", "2"], 'no value for argument "-n3"'),
])
def test_parse_errors_exit_zero_with_reference_messages(args, msg):
    r = run(*args)
    assert r.returncode == 0
    assert msg in r.stdout + r.stderr


@pytest.mark.gpu
def test_flip_graph_finds_and_saves_rank7_222(tmp_path):
    out = tmp_path / "schemes"
    print("[cli] flip_graph 2x2x2", flush=True)
    r = run("-n1", "2", "-n2", "2", "-n3", "2", "--schemes", "256", "--max-iterations", "3000",
            "--resize-probability", "0", "--max-launches", "3", "--target-rank", "7", "--seed", "1",
            "--path", str(out), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Best rank of 2x2x2 was improved from 8 to 7 (known: 7)!" in r.stdout
    assert "| run id |  size  |  real  | known | best | curr | flips count |" in r.stdout
    files = glob.glob(str(out / "*.json"))
    assert files
    for f in files:
        name = os.path.basename(f)
        assert re.fullmatch(r"2x2x2_m7_c-?\d+_iteration\d+_run\d+_2x2x2_Z2\.json", name), name
        d = json.load(open(f))
        assert d["n"] == [2, 2, 2] and d["m"] == 7 and d["z2"] is True
        blob = S.to_blob((2, 2, 2), d["u"], d["v"], d["w"])
        assert O.validate(blob) == 1
        assert d["complexity"] == O.complexity(blob)


@pytest.mark.gpu
def test_flip_graph_input_path_roundtrip(tmp_path):
    # Strassen^2 (rank 49) as --input-path text: count, then n1 n2 n3 m and U V W
    n, U, V, W = S.strassen_squared()
    blob = S.to_blob(n, U, V, W)
    txt = tmp_path / "in.txt"
    txt.write_text("1\n" + " ".join(str(int(x)) for x in blob) + "\n")
    r = run("-n1", "4", "-n2", "4", "-n3", "4", "--schemes", "128", "--max-iterations", "500",
            "--resize-probability", "0", "--max-launches", "1", "--input-path", str(txt), "--seed", "3",
            "--path", str(tmp_path / "s"), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Start reading 1 schemes" in r.stdout
    assert re.search(r"\|\s+\d+ \|  4x4x4 \|  4x4x4 \|    47 \|   49 \|", r.stdout), r.stdout[-2000:]


@pytest.mark.gpu
def test_flip_graph_resize_edges_save_valid_schemes_of_new_formats(tmp_path):
    # the reference's default --resize-probability 0.2: after each launch some
    # walkers change format (resizeKernel) and updateRanks saves every format
    # whose best rank improved (src/entities/flip_graph.cu:249-277, 296-299)
    out = tmp_path / "schemes"
    print("[cli] flip_graph 3x3x3 with resize edges", flush=True)
    r = run("-n1", "3", "-n2", "3", "-n3", "3", "--schemes", "1024", "--max-iterations", "400", "--max-launches", "4",
            "--seed", "2", "--path", str(out), timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "- resized schemes grow to at most 64 terms" in r.stdout
    assert "invalid" not in r.stdout
    improved = re.findall(r"Best rank of (\d+x\d+x\d+) was improved to (\d+) \(known: -?\d+\)!", r.stdout)
    keys = {k for k, _ in improved}
    assert len(keys - {"3x3x3"}) >= 2, improved
    files = glob.glob(str(out / "*.json"))
    assert files
    for f in files:
        d = json.load(open(f))
        n = tuple(d["n"])
        name = os.path.basename(f)
        assert re.fullmatch(r"\d+x\d+x\d+_m\d+_c-?\d+_iteration\d+_run\d+_" + "x".join(map(str, n)) + r"_Z2\.json",
                            name), name
        assert name.startswith("x".join(map(str, sorted(n))) + f"_m{d['m']}_")
        blob = S.to_blob(n, d["u"], d["v"], d["w"])
        assert O.validate(blob) == 1, name


@pytest.mark.gpu
def test_flip_graph_with_basis_change_edges(tmp_path):
    out = tmp_path / "schemes"
    print("[cli] flip_graph 2x2x2 with sandwiching and basis edges", flush=True)
    r = run("-n1", "2", "-n2", "2", "-n3", "2", "--schemes", "256", "--max-iterations", "2000",
            "--resize-probability", "0", "--sandwiching-probability", "40", "--basis-probability", "40",
            "--max-launches", "3", "--target-rank", "7", "--seed", "4", "--path", str(out), timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "  - sandwiching: 40" in r.stdout and "  - basis: 40" in r.stdout
    assert "Best rank of 2x2x2 was improved from 8 to 7" in r.stdout
    for f in glob.glob(str(out / "*.json")):
        d = json.load(open(f))
        assert O.validate(S.to_blob((2, 2, 2), d["u"], d["v"], d["w"])) == 1


@pytest.mark.gpu
def test_flip_graph_integer_ring_with_resize_and_basis_edges(tmp_path):
    # the reference's shipped build is the integer ring (src/config.cuh:6) with
    # resize edges on by default
    out = tmp_path / "schemes"
    print("[cli] flip_graph ZT 3x3x3 with resize and basis edges", flush=True)
    r = run("-n1", "3", "-n2", "3", "-n3", "3", "--ring", "ZT", "--schemes", "512", "--max-iterations", "400",
            "--basis-probability", "20", "--max-launches", "3", "--seed", "6", "--path", str(out), timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "invalid" not in r.stdout
    files = glob.glob(str(out / "*_ZT.json"))
    assert files
    for f in files:
        d = json.load(open(f))
        n = tuple(d["n"])
        assert O.validate(S.to_blob(n, d["u"], d["v"], d["w"]), O.RING_ZT) == 1, f
