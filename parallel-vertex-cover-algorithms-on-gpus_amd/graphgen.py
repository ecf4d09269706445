"""Seeded graph generators for the benchmark workloads.

Every generator draws from numpy's PCG64 bit generator seeded with the given
integer, so a (generator, parameters, seed) triple names one graph exactly;
`edge_checksum` gives the sha256 of the sorted edge list for fixtures.
Edges are 0-indexed pairs (u < v); `write_graph_file` writes the reference's
1-indexed text format (src/auxFunctions.cpp:129-153).

Workloads:
  single_er      Erdos-Renyi G(n=200, p=0.05), seed 0 (plain MVC)
  single_ba      Barabasi-Albert n=2000, m=3, seed 1 (PVC at k = opt-1, opt)
  single_reg3    3-regular configuration model n=120 with 5% of the edges
                 rewired, seed 0 (hard MVC)
  batch          4096 graphs, n ~ U[40, 400], even i Erdos-Renyi with
                 p ~ U[0.02, 0.15], odd i Barabasi-Albert with m in {2,3,4},
                 graph i from seed base+i
"""
from __future__ import annotations

import hashlib
import math

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def erdos_renyi(n: int, p: float, seed: int) -> np.ndarray:
    """Every pair i < j (row-major order) is kept when its uniform draw is < p."""
    rng = _rng(seed)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(iu.shape[0]) < p
    return np.stack([iu[keep], ju[keep]], axis=1).astype(np.uint32)


def barabasi_albert(n: int, m: int, seed: int) -> np.ndarray:
    """Preferential attachment in the networkx convention: start from a star
    on m+1 vertices (centre 0); each new vertex attaches to m distinct targets
    drawn uniformly from the list of edge endpoints (so with probability
    proportional to degree).  Edge count m + (n-m-1)*m."""
    if n <= m:
        raise ValueError("n must exceed m")
    rng = _rng(seed)
    edges = [(0, i) for i in range(1, m + 1)]
    rep = [0] * m + list(range(1, m + 1))
    for v in range(m + 1, n):
        targets: list[int] = []
        seen = set()
        while len(targets) < m:
            t = rep[int(rng.integers(len(rep)))]
            if t not in seen:
                seen.add(t)
                targets.append(t)
        for t in targets:
            edges.append((t, v))
        rep.extend(targets)
        rep.extend([v] * m)
    return np.asarray(edges, dtype=np.uint32).reshape(-1, 2)


def regular3_rewired(n: int, seed: int, rewire: float = 0.05) -> np.ndarray:
    """3-regular graph from the configuration model (uniform stub pairing,
    redrawn until there is no loop or repeated edge), then ceil(rewire*m)
    times: delete a uniformly chosen edge and add a uniformly chosen non-edge."""
    if (3 * n) % 2:
        raise ValueError("3n must be even")
    rng = _rng(seed)
    while True:
        stubs = np.repeat(np.arange(n), 3)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        es = set()
        ok = True
        for a, b in pairs.tolist():
            e = (min(a, b), max(a, b))
            if a == b or e in es:
                ok = False
                break
            es.add(e)
        if ok:
            break
    edges = sorted(es)
    for _ in range(math.ceil(rewire * len(edges))):
        i = int(rng.integers(len(edges)))
        es.discard(edges[i])
        edges.pop(i)
        while True:
            u = int(rng.integers(n))
            v = int(rng.integers(n))
            e = (min(u, v), max(u, v))
            if u != v and e not in es:
                es.add(e)
                edges.append(e)
                break
    return np.asarray(sorted(es), dtype=np.uint32).reshape(-1, 2)


def batch_graph(i: int, base_seed: int = 1000, nmin: int = 40, nmax: int = 400):
    """Graph i of the batched workload: (n, edges, description)."""
    rng = _rng(base_seed + i)
    n = int(rng.integers(nmin, nmax + 1))
    if i % 2 == 0:
        p = float(rng.uniform(0.02, 0.15))
        return n, erdos_renyi(n, p, base_seed + i), f"er n={n} p={p:.4f}"
    m = int(rng.choice([2, 3, 4]))
    return n, barabasi_albert(n, m, base_seed + i), f"ba n={n} m={m}"


def single(name: str):
    """(n, edges) of the named single-graph workload."""
    if name == "single_er":
        return 200, erdos_renyi(200, 0.05, 0)
    if name == "single_ba":
        return 2000, barabasi_albert(2000, 3, 1)
    if name == "single_reg3":
        return 120, regular3_rewired(120, 0)
    raise KeyError(name)


def edge_checksum(edges: np.ndarray) -> str:
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
    e = np.sort(e, axis=1)
    order = np.lexsort((e[:, 1], e[:, 0]))
    return hashlib.sha256(np.ascontiguousarray(e[order]).tobytes()).hexdigest()


def write_graph_file(path: str, n: int, edges: np.ndarray) -> None:
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2) + 1
    with open(path, "w") as f:
        f.write(f"{n} {len(e)}\n")
        if len(e):
            np.savetxt(f, e, fmt="%d")
