"""The edge bound of a search-tree node (csrc/vc_bound.hpp, the predicate the
kernel calls) against the mathematics it must respect: a node may be pruned
only when no cover smaller than the incumbent exists below it.

The reference reads the incumbent before the rules (src/GlobalWorkList.cuh:101)
and again before the bound (:135) and prunes at edges >= (best - d - 1)^2 + 1
(:139).  That square is only sound when the high-degree rule ran with the same
incumbent; when another worker lowered it in between, it can prune a subtree
that holds a smaller cover.  These tests restate the node body on the CPU (rule
fixpoint with threshold T, then the bound against a lower incumbent B) and
check the kernel's predicate with brute force.
"""
import ctypes
import itertools
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parallel-vertex-cover-algorithms-on-gpus_amd", "csrc")


@pytest.fixture(scope="module")
def prunes(tmp_path_factory):
    """vcg::bound_prunes compiled for the host from the kernel's header."""
    d = tmp_path_factory.mktemp("bound")
    src = d / "bound.cpp"
    src.write_text('#include "vc_bound.hpp"\n'
                   'extern "C" int prunes(unsigned e, unsigned d, unsigned T, unsigned B)\n'
                   '{ return vcg::bound_prunes(e, d, T, B) ? 1 : 0; }\n')
    so = d / "libbound.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.prunes.argtypes = [ctypes.c_uint] * 4
    lib.prunes.restype = ctypes.c_int
    return lambda e, dd, T, B: bool(lib.prunes(e, dd, T, B))


def reference_prunes(edges, d, best):
    """src/GlobalWorkList.cuh:139 with the incumbent read at :135."""
    return d >= best or edges >= (best - d - 1) ** 2 + 1


def reduce_node(n, edges, d, T):
    """Rule fixpoint of the node body with high-degree threshold T, in the
    order of src/GlobalWorkList.cuh:107-130 (leaf, triangle, high degree;
    sequential application as src/helperFunctions.cpp:49-155).  Returns the
    live edge set and the new cover size."""
    adj = {v: set() for v in range(n)}
    for u, v in edges:
        adj[u].add(v)
        adj[v].add(u)
    live = set(range(n))

    def take(v):
        nonlocal d
        live.discard(v)
        for u in adj[v]:
            adj[u].discard(v)
        adj[v] = set()
        d += 1

    changed = True
    while changed:
        changed = False
        for v in range(n):  # leaf rule
            if v in live and len(adj[v]) == 1:
                (u,) = adj[v]
                if len(adj[u]) > 1 or u < v:
                    take(u)
        for v in range(n):  # degree-two triangle rule
            if v in live and len(adj[v]) == 2:
                a, b = sorted(adj[v])
                skip = any(len(adj[x]) == 1 or (len(adj[x]) == 2 and x < v) for x in (a, b))
                if not skip and b in adj[a]:
                    take(a)
                    take(b)
                    changed = True
        for v in range(n):  # high-degree rule
            if v in live and len(adj[v]) > 0 and len(adj[v]) + d >= T:
                take(v)
                changed = True
    rest = sorted({(min(u, v), max(u, v)) for u in adj for v in adj[u]})
    return rest, d


def brute_mvc(edges):
    """Exact minimum cover of an edge set by enumeration."""
    verts = sorted({x for e in edges for x in e})
    for size in range(len(verts) + 1):
        for c in itertools.combinations(verts, size):
            s = set(c)
            if all(u in s or v in s for u, v in edges):
                return size
    return len(verts)


def test_k23_remnant_is_not_pruned_when_the_incumbent_drops(prunes):
    # K_{2,3}: hubs 0, 1 (degree 3), leaves 2, 3, 4 (degree 2); 10 vertices
    # already in the cover; the rules ran with T = d + 5; another worker then
    # lowered the incumbent to B = d + 3
    k23 = [(a, b) for a in (0, 1) for b in (2, 3, 4)]
    d0 = 10
    T, B = d0 + 5, d0 + 3
    rest, d = reduce_node(5, k23, d0, T)
    assert d == d0 and rest == sorted(k23)  # no rule fires: leaves have two hub neighbours, hubs 3 + 10 < 15
    assert brute_mvc(rest) == 2  # the two hubs: a cover of size d + 2 = 12 < B = 13
    assert reference_prunes(len(rest), d, B)  # the reference's square prunes it (6 >= 2^2 + 1)
    assert not prunes(len(rest), d, T, B)  # the kernel's bound keeps it
    # with the incumbent unchanged the two bounds agree
    assert prunes(len(rest), d, T, T) == reference_prunes(len(rest), d, T)


def test_bound_is_the_references_when_the_incumbent_is_unchanged(prunes):
    # B == T: the product (B - d - 1)(T - d - 1) is the reference's square
    for T in range(0, 40):
        for d in range(0, 45):
            for e in (0, 1, 5, 17, 80, 400, 1500):
                assert prunes(e, d, T, T) == reference_prunes(e, d, T), (e, d, T)
    # 64-bit product: budgets beyond 46340 (the reference's int square overflows)
    assert not prunes(2_000_000_000, 0, 60001, 60001)
    assert prunes(4_000_000_000, 0, 60001, 60001) == (4_000_000_000 > 60000 * 60000)


def test_bound_never_prunes_a_smaller_cover(prunes):
    """Random remnants, random cover sizes and a threshold T at or above the
    incumbent B: whenever the kernel's bound prunes, brute force confirms
    that no cover smaller than B exists below the node.  The reference's
    square, applied to the same nodes, is caught pruning a smaller cover."""
    rng = np.random.default_rng(20261203)
    checked = ref_false = 0
    for _ in range(3000):
        n = int(rng.integers(3, 10))
        p = float(rng.uniform(0.2, 0.9))
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        d0 = int(rng.integers(0, 6))
        B = d0 + int(rng.integers(1, 7))
        T = B + int(rng.integers(0, 5))
        rest, d = reduce_node(n, edges, d0, T)
        opt_below = d + brute_mvc(rest)
        if prunes(len(rest), d, T, B):
            assert opt_below >= B, (n, edges, d0, T, B)
            checked += 1
        if reference_prunes(len(rest), d, B) and opt_below < B:
            ref_false += 1
    assert checked > 500
    assert ref_false > 0  # the hazard is real on these nodes
