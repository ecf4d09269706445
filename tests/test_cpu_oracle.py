"""CPU-only checks of the oracle (the checker) and of the host side of the
product library.  No GPU is touched here.

What pins the oracle (a C restatement of src/Sequential.cpp,
src/SequentialParameterized.cpp, src/helperFunctions.cpp and the U
approximations of src/auxFunctions.cpp):
  * the reference's own known answers (test/runGraphs.py:8-30) on the graph
    files it ships (committed as tests/golden/graphs/*.npz);
  * golden vectors produced by running the reference's own CPU code
    (oracle/_ref) on seeded graphs: same U, same MVC, same node count, same
    per-rule call counts, same PVC answers;
  * brute force on tiny graphs, closed forms, and an independent exact solver
    (scipy's MILP) on mid-size graphs.
"""
import itertools
import json
import os
import re

import numpy as np
import pytest

import graphgen
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)

from golden.make_golden import load_reference_graph, make_graph, unpack_graph  # noqa: E402


def brute_mvc(n, edges):
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if len(edges) == 0:
        return 0
    masks = np.arange(1 << n, dtype=np.int64)
    ok = np.ones(1 << n, dtype=bool)
    for u, v in edges:
        ok &= (((masks >> u) & 1) | ((masks >> v) & 1)).astype(bool)
    pop = np.array([bin(x).count("1") for x in range(1 << n)])
    return int(pop[ok].min())


def milp_mvc(n, edges):
    """Independent exact MVC: 0/1 integer program solved by HiGHS."""
    from scipy.optimize import Bounds, LinearConstraint, milp
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if len(edges) == 0:
        return 0
    A = np.zeros((len(edges), n))
    A[np.arange(len(edges)), edges[:, 0]] = 1
    A[np.arange(len(edges)), edges[:, 1]] = 1
    res = milp(c=np.ones(n), constraints=LinearConstraint(A, lb=1, ub=np.inf), integrality=np.ones(n),
               bounds=Bounds(0, 1))
    assert res.success
    return int(round(res.fun))


# ---------------------------------------------------------------- known answers

KNOWN = json.load(open(os.path.join(GOLD, "known_answers.json")))


@pytest.mark.parametrize("name", sorted(k for k, v in KNOWN.items() if v["cpu_confirmed"]))
def test_oracle_reproduces_reference_known_answers(name):
    meta = KNOWN[name]
    n, e = unpack_graph(os.path.join(GOLD, "graphs", name + ".npz"))
    assert n == meta["n"] and len(e) == meta["m"]
    assert graphgen.edge_checksum(e) == meta["edge_sha256"]
    g = O.from_edges(n, e)
    assert O.check_graph(g)
    U = O.remove_max(g)
    assert U == meta["U_max"]
    r, st, cover = O.sequential(g, U)
    assert r == meta["mvc"]
    assert st["nodes"] == meta["oracle_nodes"]
    if cover is not None:
        assert O.cover_is_valid(g, cover) and int(cover.sum()) == r


def test_known_answer_files_match_reference_copies():
    """When the reference is mounted, the committed fixtures are exactly its files."""
    base = "/root/reference"
    if not os.path.isdir(base):
        pytest.skip("reference not mounted")
    for name, meta in KNOWN.items():
        g = load_reference_graph(os.path.join(base, meta["file"]), meta["index_base"])
        n, e = unpack_graph(os.path.join(GOLD, "graphs", name + ".npz"))
        g2 = O.from_edges(n, e)
        assert g.n == g2.n and g.m == g2.m
        assert np.array_equal(g.srcPtr, g2.srcPtr) and np.array_equal(g.dst, g2.dst)


# ---------------------------------------------------------------- golden vectors

VECS = json.load(open(os.path.join(GOLD, "oracle_vectors.json")))


def _fast_vecs():
    return [v for v in VECS if v["oracle"]["nodes"] < 3_000_000]


@pytest.mark.parametrize("vec", _fast_vecs(), ids=lambda v: v["key"])
def test_oracle_matches_golden_vectors(vec):
    n, e = make_graph(vec["kind"], vec["spec"])
    assert graphgen.edge_checksum(e) == vec["edge_sha256"]
    g = O.from_edges(n, e)
    assert O.remove_max(g) == vec["U_max"]
    assert O.remove_edge(g, vec["U_edge_seed"]) == vec["U_edge"]
    r, st, _ = O.sequential(g, vec["U_max"])
    assert r == vec["mvc"]
    for f, want in vec["oracle"].items():
        assert st[f] == want, f
    for k, want in vec["pvc"].items():
        kf, pst, _ = O.sequential_parameterized(g, vec["U_max"], int(k))
        assert kf == want["kFound"] and pst["nodes"] == want["nodes"]


def test_golden_vectors_were_checked_against_reference_code():
    assert VECS and all(v["ref_checked"] for v in VECS)


def test_pvc_oracle_agrees_with_mvc_threshold():
    """For k < U the reference's SequentialParameterized answers [k >= MVC]
    on every golden graph."""
    for vec in VECS:
        for k, want in vec["pvc"].items():
            if int(k) < vec["U_max"]:
                assert want["kFound"] == int(int(k) >= vec["mvc"]), (vec["key"], k)


def test_reference_pvc_defect_when_k_reaches_U():
    """Reference defect, reproduced by the compiled reference itself: the CPU
    PVC passes U (not k+1) to the high-degree rule
    (src/SequentialParameterized.cpp:45), and that rule is only sound for
    covers smaller than U.  With k >= U it can answer "no" although U <= k is
    a cover size.  On the 4-cycle U = MVC = 2: at k = 2 every vertex has
    deg + 0 >= 2 and is deleted, d = 4 > k, and the root is pruned.  The
    reference's GPU kernels use k+1 (src/GlobalWorkListParameterized.cuh:127)
    and answer "yes", as this build does."""
    n, e = 4, [(0, 1), (1, 2), (2, 3), (3, 0)]
    g = O.from_edges(n, e)
    U = O.remove_max(g)
    assert U == 2 and brute_mvc(n, e) == 2
    kf, _, _ = O.sequential_parameterized(g, U, 2)
    assert kf == 0  # the reference's (wrong) answer
    if O.ref_available():
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "c4.txt")
            graphgen.write_graph_file(path, n, e)
            assert O.ref_run(path, pvc_k=2)["kFound"] == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_oracle_matches_reference_build_live():
    for s in range(4):
        n, e = 70 + 10 * s, graphgen.erdos_renyi(70 + 10 * s, 0.07, 900 + s)
        g = O.from_edges(n, e)
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "g.txt")
            graphgen.write_graph_file(path, n, e)
            ref = O.ref_run(path, seed=77)
            gl = O.load(path)
            assert np.array_equal(gl.dst, g.dst)
        U = O.remove_max(g)
        r, st, _ = O.sequential(g, U)
        assert (ref["U_max"], ref["U_edge"], ref["result"], ref["nodes"]) == (U, O.remove_edge(g, 77), r, st["nodes"])


# ---------------------------------------------------------------- independent exactness

def _tiny_cases():
    rng = np.random.default_rng(3)
    out = []
    for i in range(60):
        n = int(rng.integers(1, 15))
        out.append((n, graphgen.erdos_renyi(n, float(rng.uniform(0.1, 0.8)), 4000 + i)))
    out.append((10, [(i, (i + 1) % 5) for i in range(5)] + [(i, i + 5) for i in range(5)] +
                [(5 + i, 5 + (i + 2) % 5) for i in range(5)]))
    return out


def test_oracle_matches_bruteforce_tiny():
    for n, e in _tiny_cases():
        g = O.from_edges(n, e)
        U = O.remove_max(g)
        r, _, cover = O.sequential(g, U)
        exact = brute_mvc(n, e)
        assert U >= exact
        assert r == exact, (n, e)
        if cover is not None:
            assert O.cover_is_valid(g, cover) and int(cover.sum()) == exact
        for k in (exact - 1, exact, exact + 1):
            if k >= 0:
                kf, _, pc = O.sequential_parameterized(g, U, k)
                if k < U:
                    assert kf == int(k >= exact)
                else:
                    # k >= U: the reference can answer "no" wrongly (see
                    # test_reference_pvc_defect_when_k_reaches_U); a "yes" is right
                    assert kf == 0 or exact <= k
                if kf:
                    assert O.cover_is_valid(g, pc) and int(pc.sum()) <= k


def test_oracle_closed_forms():
    cases = [
        (lambda n: [(i, i + 1) for i in range(n - 1)], lambda n: n // 2),        # path P_n
        (lambda n: [(i, (i + 1) % n) for i in range(n)], lambda n: (n + 1) // 2),  # cycle C_n
        (lambda n: list(itertools.combinations(range(n), 2)), lambda n: n - 1),  # complete K_n
        (lambda n: [(0, i) for i in range(1, n)], lambda n: 1),                  # star
    ]
    for mk, f in cases:
        for n in (3, 4, 7, 12, 25):
            g = O.from_edges(n, mk(n))
            r, _, _ = O.sequential(g, O.remove_max(g))
            assert r == f(n)
    for a, b in ((2, 5), (4, 4), (6, 9)):
        g = O.from_edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])
        assert O.sequential(g, O.remove_max(g))[0] == min(a, b)   # Konig on K_{a,b}


def test_oracle_matches_milp_midsize():
    for s in range(6):
        n = 45 + 5 * s
        e = graphgen.erdos_renyi(n, 0.08, 600 + s)
        g = O.from_edges(n, e)
        assert O.sequential(g, O.remove_max(g))[0] == milp_mvc(n, e)
    n, e = 150, graphgen.barabasi_albert(150, 3, 5)
    g = O.from_edges(n, e)
    assert O.sequential(g, O.remove_max(g))[0] == milp_mvc(n, e)


PLANTS = {
    1: "edge bound (U-d)^2 instead of (U-d-1)^2 (src/Sequential.cpp:67)",
    2: "branch on the highest-index max-degree vertex (src/Sequential.cpp:50-58)",
    3: "triangle rule deleting only the first neighbour (src/helperFunctions.cpp:144-147)",
    4: "high-degree rule with > instead of >= (src/helperFunctions.cpp:85)",
    5: "RemoveMax deleting the last max-degree vertex (src/auxFunctions.cpp:236-243)",
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_oracle_mistakes_are_caught(plant, tmp_path):
    """Build the oracle with one plausible mistake planted (-DORACLE_PLANT)
    and check that the golden vectors from the reference's own code reject it."""
    import subprocess
    so = tmp_path / f"liboracle_plant{plant}.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-pthread", f"-DORACLE_PLANT={plant}",
                    os.path.join(ROOT, "oracle", "vc_oracle.c"), "-o", str(so)], check=True)
    mutant = O.load_library(str(so))
    saved = O._lib
    O._lib = mutant
    caught = False
    try:
        for vec in [v for v in VECS if v["oracle"]["nodes"] < 200_000]:
            n, e = make_graph(vec["kind"], vec["spec"])
            g = O.from_edges(n, e)
            U = O.remove_max(g)
            r, st, _ = O.sequential(g, vec["U_max"])
            got = (U, r, st["nodes"], st["leaf_calls"], st["tri_calls"], st["high_calls"], st["delete_calls"])
            want = (vec["U_max"], vec["mvc"], vec["oracle"]["nodes"], vec["oracle"]["leaf_calls"],
                    vec["oracle"]["tri_calls"], vec["oracle"]["high_calls"], vec["oracle"]["delete_calls"])
            if got != want:
                caught = True
                break
    finally:
        O._lib = saved
    assert caught, f"planted mistake not detected: {PLANTS[plant]}"


def test_node_budget_stops_search():
    vec = next(v for v in VECS if v["oracle"]["nodes"] > 1000)
    n, e = make_graph(vec["kind"], vec["spec"])
    g = O.from_edges(n, e)
    r, st, _ = O.sequential(g, vec["U_max"], max_nodes=100)
    assert st["nodes"] == 100 and st["completed"] == 0


# ---------------------------------------------------------------- loader / approximations of the product (host side)

def _vc():
    import vcgpu
    try:
        vcgpu.lib()
    except RuntimeError as ex:
        pytest.skip(str(ex))
    return vcgpu


def test_c_abi_exports_every_declared_symbol():
    vcgpu = _vc()
    header = open(os.path.join(ROOT, "include", "vcgpu.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vc_[a-z_0-9]+)\s*\(", header, flags=re.M)))
    assert declared, "no declarations parsed"
    L = vcgpu.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(vcgpu.EXPORTS) == declared


def test_product_loader_matches_oracle_loader(tmp_path):
    vcgpu = _vc()
    for s in range(3):
        n, e = 80, graphgen.erdos_renyi(80, 0.1, 70 + s)
        p = tmp_path / f"g{s}.txt"
        graphgen.write_graph_file(str(p), n, e)
        a = vcgpu.load_graph_file(str(p))
        b = O.load(str(p))
        assert a.n == b.n and a.m == b.m
        assert np.array_equal(a.srcPtr, b.srcPtr) and np.array_equal(a.dst, b.dst)
        assert np.array_equal(a.degree, b.degree)
        assert vcgpu.check_graph(a)


def test_product_rejects_bad_graphs(tmp_path):
    vcgpu = _vc()
    p = tmp_path / "dup.txt"
    p.write_text("3 2\n1 2\n1 2\n")
    g = vcgpu.load_graph_file(str(p))
    assert not vcgpu.check_graph(g)
    p2 = tmp_path / "loop.txt"
    p2.write_text("3 1\n2 2\n")
    assert not vcgpu.check_graph(vcgpu.load_graph_file(str(p2)))
    p3 = tmp_path / "range.txt"
    p3.write_text("3 1\n1 4\n")
    with pytest.raises(vcgpu.VCError):
        vcgpu.load_graph_file(str(p3))


def test_product_check_graph_matches_reference_semantics():
    """vc_check_graph (a last-seen marker, O(n + m)) answers exactly as the
    reference's n x n matrix check, restated in the oracle, on adjacency
    lists with duplicates in any order and self-loops at random places."""
    vcgpu = _vc()
    rng = np.random.Generator(np.random.PCG64(5))
    outcomes = set()
    for _ in range(400):
        n = int(rng.integers(1, 25))
        lists = []
        for v in range(n):
            k = int(rng.integers(0, min(n, 7)))
            if rng.random() < 0.5:
                others = [u for u in range(n) if u != v]
                nb = list(rng.choice(others, size=min(k, len(others)), replace=False)) if others else []
            else:
                nb = list(rng.integers(0, n, size=k))
            rng.shuffle(nb)
            lists.append([int(u) for u in nb])
        deg = np.array([len(x) for x in lists], np.int32)
        src = np.zeros(n + 1, np.uint32)
        src[1:] = np.cumsum(deg)
        total = int(src[-1])
        m = (total + 1) // 2
        dst = np.zeros(max(2 * m, 1), np.uint32)
        dst[:total] = [u for x in lists for u in x]
        # the definition: no list repeats a vertex or names its own vertex
        brute = all(len(set(x)) == len(x) and v not in x for v, x in enumerate(lists))
        want = O.check_graph(O.Graph(n, m, src, dst, deg))
        got = vcgpu.check_graph(vcgpu.CSR(n, m, src, dst, deg))
        assert want == brute and got == brute, lists
        outcomes.add(brute)
    assert outcomes == {True, False}


def test_product_loader_reads_what_the_reference_reads(tmp_path):
    """Tabs, CR LF line ends, blank and trailing lines and a header with a
    trailing space (the .mtx files) parse as the reference's fscanf loop
    parses them; a file with fewer edges than its header says is an error."""
    vcgpu = _vc()
    n, e = 60, graphgen.erdos_renyi(60, 0.1, 11)
    lines = [f"{n} {len(e)} "] + [f"{u + 1}\t{v + 1}" for u, v in e]
    p = tmp_path / "odd.txt"
    p.write_bytes(("\r\n".join(lines) + "\r\n\n\n").encode())
    a = vcgpu.load_graph_file(str(p))
    b = O.load(str(p))
    assert (a.n, a.m) == (b.n, b.m) == (n, len(e))
    assert np.array_equal(a.srcPtr, b.srcPtr) and np.array_equal(a.dst, b.dst)
    short = tmp_path / "short.txt"
    short.write_text("\n".join([f"{n} {len(e) + 3}"] + [f"{u + 1} {v + 1}" for u, v in e]) + "\n")
    with pytest.raises(vcgpu.VCError):
        vcgpu.load_graph_file(str(short))


@pytest.mark.parametrize("name", sorted(k for k, v in KNOWN.items() if v["index_base"] == 0))
def test_product_loads_zero_indexed_reference_files(name, tmp_path):
    """vc_load_graph_file_indexed(path, 0, ...) reads the reference's
    0-indexed files (here rewritten from the fixture in their layout: header,
    then one edge a line, tab-separated) into the CSR the oracle builds from
    the same edges; the 1-indexed loader refuses them (vertex 0)."""
    vcgpu = _vc()
    meta = KNOWN[name]
    n, e = unpack_graph(os.path.join(GOLD, "graphs", name + ".npz"))
    p = tmp_path / (name + ".txt")
    p.write_text(f"{n} {len(e)}\n" + "".join(f"{u}\t{v}\n" for u, v in e))
    a = vcgpu.load_graph_file(str(p), index_base=0)
    b = O.from_edges(n, e)
    assert (a.n, a.m) == (meta["n"], meta["m"])
    assert np.array_equal(a.srcPtr, b.srcPtr) and np.array_equal(a.dst, b.dst)
    assert np.array_equal(a.degree, b.degree) and vcgpu.check_graph(a)
    if int(e.min()) == 0:
        with pytest.raises(vcgpu.VCError):
            vcgpu.load_graph_file(str(p))


def test_product_index_base_errors(tmp_path):
    vcgpu = _vc()
    p = tmp_path / "z.txt"
    p.write_text("3 2\n0 1\n1 2\n")
    g = vcgpu.load_graph_file(str(p), index_base=0)
    assert (g.n, g.m) == (3, 2) and list(g.degree) == [1, 2, 1]
    p.write_text("3 2\n0 1\n1 3\n")  # vertex 3 is out of range when counting from 0
    with pytest.raises(vcgpu.VCError):
        vcgpu.load_graph_file(str(p), index_base=0)
    with pytest.raises(vcgpu.VCError):
        vcgpu.load_graph_file(str(p), index_base=2)


@pytest.mark.parametrize("vec", VECS[:40], ids=lambda v: v["key"])
def test_product_upper_bounds_match_reference(vec):
    """vc_approx_remove_max / vc_approx_remove_edge give the reference's U
    bit for bit (golden vectors from oracle/_ref), and their covers are valid."""
    vcgpu = _vc()
    n, e = make_graph(vec["kind"], vec["spec"])
    g = vcgpu.csr_from_edges(n, e)
    og = O.from_edges(n, e)
    u, cov = vcgpu.approx_remove_max(g)
    assert u == vec["U_max"]
    assert O.cover_is_valid(og, cov) and int(cov.sum()) == u
    ue, cove = vcgpu.approx_remove_edge(g, vec["U_edge_seed"])
    assert ue == vec["U_edge"]
    assert O.cover_is_valid(og, cove) and int(cove.sum()) == ue


def _bound_cases():
    """Graph families whose rule passes exercise the candidate sets of the
    product's approximations: leaves in cascades (paths, trees, stars),
    degree-2 chains and triangles (cycles, triangle chains, 3-regular),
    dense and sparse random graphs, disjoint unions, one BA n = 20,000."""
    rng = np.random.default_rng(4242)
    for n in (1, 2, 3, 5, 13, 60):
        yield n, [(i, i + 1) for i in range(n - 1)]
        if n >= 3:
            yield n, [(i, (i + 1) % n) for i in range(n)]
        yield n, [(0, i) for i in range(1, n)]
    yield 12, [(i, j) for i in range(12) for j in range(i + 1, 12)]
    for k in (1, 9):
        yield 2 * k + 1, [e for i in range(k) for e in ((2 * i, 2 * i + 1), (2 * i + 1, 2 * i + 2), (2 * i, 2 * i + 2))]
    for _ in range(15):
        n = int(rng.integers(2, 300))
        yield n, [(int(rng.integers(0, i)), i) for i in range(1, n)]
    for _ in range(40):
        n = int(rng.integers(5, 400))
        yield n, graphgen.erdos_renyi(n, float(rng.uniform(0.005, 0.2)), int(rng.integers(1 << 30)))
    for _ in range(20):
        n, m = int(rng.integers(8, 800)), int(rng.integers(1, 6))
        yield n, graphgen.barabasi_albert(n, m, int(rng.integers(1 << 30)))
    for _ in range(8):
        yield 120, graphgen.regular3_rewired(120, int(rng.integers(1 << 30)))
    for _ in range(8):
        off, edges = 0, []
        for _ in range(int(rng.integers(2, 10))):
            n = int(rng.integers(2, 30))
            e = np.asarray(graphgen.erdos_renyi(n, float(rng.uniform(0.1, 0.6)), int(rng.integers(1 << 30))))
            edges += [(off + int(a), off + int(b)) for a, b in e.reshape(-1, 2)]
            off += n
        yield off, edges
    yield 20000, graphgen.barabasi_albert(20000, 4, 5000)


def test_product_upper_bounds_match_oracle_on_many_graphs():
    """The product's candidate-driven approximations delete the same
    vertices as the reference's full-scan passes: same U as the oracle's
    restatement (which scans like the reference) for RemoveMax and for
    RemoveEdge at several seeds, on every family above, with valid covers."""
    vcgpu = _vc()
    count = 0
    for n, e in _bound_cases():
        e = np.asarray(e, dtype=np.uint32).reshape(-1, 2)
        g, og = vcgpu.csr_from_edges(n, e), O.from_edges(n, e)
        u, cov = vcgpu.approx_remove_max(g)
        assert u == O.remove_max(og), (n, len(e))
        assert O.cover_is_valid(og, cov) and int(cov.sum()) == u
        for seed in (0, 7, 1000):
            ue, cove = vcgpu.approx_remove_edge(g, seed)
            assert ue == O.remove_edge(og, seed), (n, len(e), seed)
            assert O.cover_is_valid(og, cove) and int(cove.sum()) == ue
        count += 1
    assert count == 111


def test_generator_checksums_are_stable():
    want = {v["spec"]["name"]: v["edge_sha256"] for v in VECS if v["kind"] == "single"}
    for name, cs in want.items():
        n, e = graphgen.single(name)
        assert graphgen.edge_checksum(e) == cs
    n, e = graphgen.single("single_ba")
    assert n == 2000 and len(e) == 3 + 1996 * 3
    n, e = graphgen.single("single_reg3")
    assert n == 120 and len(e) == 180
    deg = np.bincount(np.asarray(e).ravel(), minlength=n)
    assert deg.sum() == 360
