"""The served mixed-stream workload: Poisson arrivals of MVC and PVC requests
through continuous-batching servers (vc_server_*), shared by bench.py and
the GPU tests.

Stream (one request = one graph + one job):
  small class (95%)  graph i of the batched distribution (graphgen.batch_graph:
                     n ~ U[40, 400], half Erdos-Renyi p ~ U[0.02, 0.15], half
                     Barabasi-Albert m in {2,3,4}), i uniform over the pool.
                     Its "alone" answer comes from an offline batch solve of
                     the pool with the same node budget.  Graphs whose MVC
                     that solve established get, 50/50, an MVC request
                     (expected: that MVC) or a PVC request at k = opt - 1 or
                     opt + 1 (expected: [k >= opt]).  Graphs whose MVC it did
                     not establish get a budgeted MVC request (expected: an
                     incumbent <= U; reported unsolved when the budget ends
                     it, exact if the search finishes).
  large class (5%)   Barabasi-Albert n = 20,000, m = 4.  Its MVC is out of
                     reach of this rule set (U ~ 10,800, no degree-1 or
                     degree-2 vertex at the root), so opt, and with it
                     k = opt +- 1, is unknown.  Stand-ins: 50/50 a budgeted
                     MVC request (expected: an incumbent <= U whose cover
                     checks; never counted exact) or PVC at k = U (expected:
                     yes, certified only by the device cover the request
                     returns, which must check and have size <= k; reported
                     apart from exact answers, since the approximation's own
                     cover already shows the answer).

Two servers share the GPU: the large class on one whose node state lives in
HBM (started first: a persistent kernel launched second only gets the CUs
left over), the small class on one with LDS node state.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np

import graphgen
import vcgpu

LARGE_N, LARGE_M = 20000, 4


@dataclass
class Request:
    gi: int             # index into Stream.graphs
    job: vcgpu.Job
    cls: str            # "small" | "large"
    kind: str           # "mvc" | "pvc" | "mvc_budget" | "large_mvc" | "large_pvc_at_U"
    expect: int | None  # MVC size / PVC answer when known


@dataclass
class Stream:
    graphs: list        # vcgpu.CSR
    edges: list         # np.ndarray (m, 2) per graph
    ubs: list           # (U, cover) per graph
    alone: list         # vcgpu.Result of the offline batch solve (small graphs), None for large
    requests: list      # Request
    gaps: np.ndarray    # inter-arrival times in seconds
    precompute_s: float


def build_stream(ctx, n_req: int, rate: float, base_seed: int = 1000, pool: int = 4096, large_graphs: int = 8,
                 small_budget: int = 1 << 20, large_budget: int = 1 << 12, large_share: float = 0.05,
                 time_limit_ms: int = 120000, seed: int = 77, pool_graphs=None, pool_ubs=None,
                 pool_alone=None) -> Stream:
    """pool_graphs / pool_ubs / pool_alone: the batch graphs 0..pool-1 with
    their bounds and the results of a batch solve of them (with at least the
    request budget), when the caller has them already."""
    rng = np.random.Generator(np.random.PCG64(base_seed + seed))
    small = [graphgen.batch_graph(i, base_seed)[:2] for i in range(pool)]
    large = [(LARGE_N, graphgen.barabasi_albert(LARGE_N, LARGE_M, 5000 + j)) for j in range(large_graphs)]
    edges = [np.asarray(e, np.uint32).reshape(-1, 2) for _, e in small + large]
    if pool_graphs is not None:
        graphs = list(pool_graphs[:pool]) + [vcgpu.csr_from_edges(n, e) for (n, _), e in zip(large, edges[pool:])]
        ubs = list(pool_ubs[:pool]) + [vcgpu.upper_bound(g, seed=base_seed) for g in graphs[pool:]]
    else:
        graphs = [vcgpu.csr_from_edges(n, e) for (n, _), e in zip(small + large, edges)]
        ubs = [vcgpu.upper_bound(g, seed=base_seed) for g in graphs]
    t0 = time.perf_counter()
    if pool_alone is not None:
        alone = list(pool_alone[:pool])
    else:
        gs = ctx.upload(graphs[:pool])
        alone, _, _ = ctx.solve(gs, [vcgpu.Job(vcgpu.VC_MVC, 0, u, small_budget) for u, _ in ubs[:pool]],
                                vcgpu.Params(time_limit_ms=time_limit_ms), want_covers=False, raise_on_error=False)
        gs.free()
    precompute_s = time.perf_counter() - t0
    reqs = []
    for _ in range(n_req):
        big = rng.random() < large_share
        mvc = rng.random() < 0.5
        if big:
            gi = pool + int(rng.integers(large_graphs))
            U = ubs[gi][0]
            if mvc:
                reqs.append(Request(gi, vcgpu.Job(vcgpu.VC_MVC, 0, U, large_budget), "large", "large_mvc", None))
            else:
                reqs.append(Request(gi, vcgpu.Job(vcgpu.VC_PVC, U, U, large_budget), "large", "large_pvc_at_U", 1))
            continue
        gi = int(rng.integers(pool))
        U = ubs[gi][0]
        r = alone[gi]
        if not r.solved:
            reqs.append(Request(gi, vcgpu.Job(vcgpu.VC_MVC, 0, U, small_budget), "small", "mvc_budget", None))
        elif mvc:
            reqs.append(Request(gi, vcgpu.Job(vcgpu.VC_MVC, 0, U, small_budget), "small", "mvc", r.minimum))
        else:
            k = max(0, r.minimum + (1 if rng.random() < 0.5 else -1))
            reqs.append(Request(gi, vcgpu.Job(vcgpu.VC_PVC, k, U, small_budget), "small", "pvc",
                                int(k >= r.minimum)))
    gaps = rng.exponential(1.0 / rate, size=n_req)
    return Stream(graphs, edges, ubs, list(alone) + [None] * large_graphs, reqs, gaps, precompute_s)


def run_stream(ctx, st: Stream, small_blocks: int = 0, large_blocks: int = 32, time_limit_ms: int = 120000,
               want_large_covers: bool = True):
    """Submits every request at its Poisson arrival time and collects the
    results.  Returns (results, arrivals_ns, start_ns, small_info, large_info,
    large_stack_depth); results[r] = (vcgpu.Result, done_ns)."""
    large_ids = sorted({q.gi for q in st.requests if q.cls == "large"})
    nmax_small = max([st.graphs[q.gi].n for q in st.requests if q.cls == "small"] + [1])
    large_srv = None
    large_depth = 0
    if large_ids:
        # the large server's stacks hold a full path of any large request
        # (admission refuses requests that could need more)
        large_depth = max(min(st.graphs[gi].n, st.ubs[gi][0] + 1) + 1 for gi in large_ids)
        large_srv = vcgpu.Server(ctx, max_vertices=max(st.graphs[gi].n for gi in large_ids), slots=64,
                                 arena_bytes=256 << 20, block_dim=64, num_blocks=large_blocks,
                                 stack_depth=large_depth, time_limit_ms=time_limit_ms)
    small_srv = vcgpu.Server(ctx, max_vertices=nmax_small, slots=4096, arena_bytes=256 << 20,
                             num_blocks=small_blocks, time_limit_ms=time_limit_ms)
    n_req = len(st.requests)
    arrivals = [0] * n_req
    tickets = [None] * n_req
    results = [None] * n_req
    errors = []
    aborted = threading.Event()

    def collect():
        try:
            for r in range(n_req):
                while tickets[r] is None:
                    if aborted.is_set():
                        return
                    time.sleep(0.0002)
                srv, t = tickets[r]
                results[r] = srv.wait(t, time_limit_ms,
                                      want_cover=want_large_covers or st.requests[r].cls == "small")
        except Exception as exc:  # surfaced to the caller after the join
            errors.append(exc)

    th = threading.Thread(target=collect)
    th.start()
    start = vcgpu.now_ns()
    due = start
    try:
        for r, q in enumerate(st.requests):
            due += int(st.gaps[r] * 1e9)
            while True:
                left = due - vcgpu.now_ns()
                if left <= 0:
                    break
                if left > 300_000:
                    time.sleep((left - 200_000) / 1e9)
            srv = large_srv if q.cls == "large" else small_srv
            arrivals[r] = vcgpu.now_ns()
            tickets[r] = (srv, srv.submit(st.graphs[q.gi], q.job))
    except BaseException:
        aborted.set()
        raise
    finally:
        th.join()
        info_s = small_srv.stop()
        info_l = large_srv.stop() if large_srv else None
    if errors:
        raise errors[0]
    return results, arrivals, start, info_s, info_l, large_depth


def cover_checks(edges: np.ndarray, cover, k: int | None = None) -> bool:
    """cover (bool per vertex) covers every edge, and has at most k vertices when k is given."""
    if cover is None:
        return False
    c = np.asarray(cover, bool)
    ok = bool(np.all(c[edges[:, 0]] | c[edges[:, 1]])) if len(edges) else True
    return ok and (k is None or int(c.sum()) <= k)


def score(st: Stream, results, arrivals, start_ns):
    """Per-class counts and latencies of a served run.

    exact        finished searches whose answer equals the expected one (the
                 alone answer: MVC size / PVC answer), or a finished
                 budgeted MVC search (no expected value)
    unsolved     searches the node budget ended (an MVC incumbent still
                 checks: <= U, a valid cover of that size)
    trivial_yes  large PVC at k = U answered yes with a checking cover
    improved     large MVC incumbents below U
    mismatches   a wrong answer, a cover that does not check or exceeds k,
                 or an MVC incumbent above U
    """
    out = {"small": dict(n=0, exact=0, unsolved=0, mismatches=0, lat=[]),
           "large": dict(n=0, exact=0, unsolved=0, trivial_yes=0, mismatches=0, improved=0, lat=[])}
    last = start_ns
    bad = []
    for r, q in enumerate(st.requests):
        c = out[q.cls]
        c["n"] += 1
        if results[r] is None:
            c["mismatches"] += 1
            bad.append((r, "no result"))
            continue
        res, done_ns = results[r]
        c["lat"].append((done_ns - arrivals[r]) / 1e6)
        last = max(last, done_ns)
        e = st.edges[q.gi]
        U, ucov = st.ubs[q.gi]
        if q.job.instance == vcgpu.VC_MVC:
            cover = res.cover if res.cover is not None else ucov
            ok = res.minimum <= U and cover_checks(e, cover) and int(np.sum(cover)) == res.minimum
            if ok and res.solved:
                ok = q.expect is None or res.minimum == q.expect
                c["exact"] += int(ok)
            elif ok:
                c["unsolved"] += 1
            if ok and q.cls == "large" and res.minimum < U:
                c["improved"] += 1
        else:
            ok = not res.k_found or cover_checks(e, res.cover, q.job.k)
            if ok and res.solved:
                ok = res.k_found == q.expect
                if ok:
                    c["trivial_yes" if q.kind == "large_pvc_at_U" else "exact"] += 1
            elif ok:
                c["unsolved"] += 1
        if not ok:
            c["mismatches"] += 1
            bad.append((r, q.kind, res))
    wall = max(last - start_ns, 1) / 1e9
    return out, wall, bad
