"""The drop-in command line (`vc`, csrc/vc_main.cpp): the reference's flags
(src/config.h:178-243), its stdout lines and its Results/Results.csv row
(src/auxFunctions.cpp:580-614)."""
import csv
import os
import re
import subprocess

import numpy as np
import pytest

import graphgen
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
VC = os.path.join(os.path.dirname(HERE), "parallel-vertex-cover-algorithms-on-gpus_amd", "vc")


def run_vc(args, cwd, seed=7):
    env = dict(os.environ, VC_SEED=str(seed))
    return subprocess.run([VC] + args, cwd=str(cwd), capture_output=True, text=True, timeout=120, env=env)


def _graph(tmp_path, n=80, p=0.1, seed=3, name="g.txt"):
    e = graphgen.erdos_renyi(n, p, seed)
    path = tmp_path / name
    graphgen.write_graph_file(str(path), n, e)
    return path, n, e


def test_cli_upper_bounds_are_the_references(tmp_path):
    """The two approximation lines carry the reference's U values (the oracle
    restates RemoveMax / RemoveEdge bit for bit); -v sequential, the
    reference's CPU solver, is refused after them."""
    path, n, e = _graph(tmp_path)
    r = run_vc(["-f", path.name, "-v", "sequential"], tmp_path, seed=7)
    assert r.returncode == 2 and "not part of this GPU build" in r.stderr
    g = O.from_edges(n, e)
    got_max = int(re.search(r"Approximate Remove Max Minimum is: (\d+)", r.stdout).group(1))
    got_edge = int(re.search(r"Approximate Remove Edge Minimum is: (\d+)", r.stdout).group(1))
    assert got_max == O.remove_max(g)
    assert got_edge == O.remove_edge(g, 7)
    assert "Graph file: g.txt" in r.stdout


def _zero_indexed_fixture(tmp_path, name):
    """A 0-indexed reference graph, written back in its file layout from the
    committed fixture."""
    from golden.make_golden import unpack_graph
    import json
    meta = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))[name]
    n, e = unpack_graph(os.path.join(HERE, "golden", "graphs", name + ".npz"))
    path = tmp_path / (name + ".txt")
    path.write_text(f"{n} {len(e)}\n" + "".join(f"{u}\t{v}\n" for u, v in e))
    return path, n, e, meta


def test_cli_reads_zero_indexed_files_with_z(tmp_path):
    """-z reads a file whose vertices start at 0 (the reference ships six with
    answers in test/runGraphs.py:21-29); without it the file is refused."""
    path, n, e, meta = _zero_indexed_fixture(tmp_path, "formated_maayan-figeys")
    r = run_vc(["-f", path.name, "-z", "-v", "sequential"], tmp_path)
    assert r.returncode == 2 and "not part of this GPU build" in r.stderr
    got_max = int(re.search(r"Approximate Remove Max Minimum is: (\d+)", r.stdout).group(1))
    assert got_max == O.remove_max(O.from_edges(n, e))
    r = run_vc(["-f", path.name], tmp_path)
    assert r.returncode == 2 and "cannot load" in r.stderr


def test_cli_rejects_bad_input(tmp_path):
    path, _, _ = _graph(tmp_path)
    assert "Unrecognized option" in run_vc(["-x"], tmp_path).stderr
    assert run_vc(["-f", path.name, "-v", "fastest"], tmp_path).returncode == 2
    r = run_vc(["-f", path.name, "-q", "1000"], tmp_path)
    assert r.returncode == 2 and "power of 2" in r.stderr
    dup = tmp_path / "dup.txt"
    dup.write_text("3 2\n1 2\n2 1\n")
    r = run_vc(["-f", dup.name], tmp_path)
    assert r.returncode == 2 and "graph check failed" in r.stderr
    assert run_vc(["-f", "missing.txt"], tmp_path).returncode == 2


@pytest.mark.gpu
def test_cli_solves_and_appends_the_results_row(tmp_path):
    """MVC and PVC through the CLI: the reference's answer lines, and one
    22-column Results.csv row per run (src/auxFunctions.cpp:589-593)."""
    path, n, e = _graph(tmp_path, n=70, p=0.08, seed=21)
    opt, _, _ = O.mvc(O.from_edges(n, e))
    r = run_vc(["-f", path.name, "-o", "run1"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert int(re.search(r"Size of minimum vertex cover: (\d+)", r.stdout).group(1)) == opt
    for k, word in ((opt - 1, "greater than"), (opt, "less than or equal to")):
        r = run_vc(["-f", path.name, "-i", "pvc", "-k", str(k), "-v", "stackonly", "-o", f"pvc{k}"], tmp_path)
        assert r.returncode == 0, r.stderr
        assert f"Minimum is {word} K: {k}" in r.stdout
    rows = list(csv.reader(open(tmp_path / "Results" / "Results.csv")))
    assert len(rows) == 3 and all(len(row) == 22 for row in rows)
    assert rows[0][0] == "run1" and rows[0][4] == "mvc" and int(rows[0][18]) == opt
    assert rows[1][4] == "pvc" and rows[1][5] == "stackonly" and int(rows[1][20]) == 0
    assert int(rows[2][20]) == 1


@pytest.mark.gpu
def test_cli_solves_a_zero_indexed_reference_file(tmp_path):
    """The reference's MVC for formated_maayan-vidal (test/runGraphs.py:28)
    through `vc -z`."""
    path, n, e, meta = _zero_indexed_fixture(tmp_path, "formated_maayan-vidal")
    r = run_vc(["-f", path.name, "-z"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert int(re.search(r"Size of minimum vertex cover: (\d+)", r.stdout).group(1)) == meta["mvc"] == 1077


@pytest.mark.gpu
@pytest.mark.parametrize("block_dim", [512, 1024])
def test_cli_block_dims_of_the_reference_sweep(tmp_path, block_dim):
    """-b 512 and -b 1024, the top of the reference harness's BLOCK_DIM sweep
    (test/runGraphs.py:7; accepted up to 1024 by src/auxFunctions.cpp:409-444):
    8 or 16 worker waves per workgroup, the same MVC, PVC answers and CSV row."""
    path, n, e = _graph(tmp_path, n=70, p=0.08, seed=21)
    opt, _, _ = O.mvc(O.from_edges(n, e))
    r = run_vc(["-f", path.name, "-b", str(block_dim), "-o", "b"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert int(re.search(r"Size of minimum vertex cover: (\d+)", r.stdout).group(1)) == opt
    assert f"NumOfThreadPerBlocks : {block_dim}" in r.stdout
    for k, word in ((opt - 1, "greater than"), (opt, "less than or equal to")):
        r = run_vc(["-f", path.name, "-i", "pvc", "-k", str(k), "-b", str(block_dim), "-g", "1"], tmp_path)
        assert r.returncode == 0, r.stderr
        assert f"Minimum is {word} K: {k}" in r.stdout
    rows = list(csv.reader(open(tmp_path / "Results" / "Results.csv")))
    assert len(rows) == 3 and all(len(row) == 22 for row in rows)
    assert all(int(row[10]) == block_dim for row in rows)


def test_cli_rejects_a_workgroup_above_1024(tmp_path):
    """Workgroups beyond the kernels' 1024-thread launch bound, or not a whole
    number of 64-lane waves, are refused with an error line (no device
    work is started for them)."""
    path, _, _ = _graph(tmp_path)
    for b in ("2048", "96"):
        r = run_vc(["-f", path.name, "-b", b], tmp_path)
        assert r.returncode == 2 and "-b must be a multiple of 64" in r.stderr
