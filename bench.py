"""Benchmark of the gfx950 vertex-cover solver on one MI355X.

Default workload (batched MVC): 4096 graphs, n ~ U[40, 400], even i
Erdos-Renyi with p ~ U[0.02, 0.15], odd i Barabasi-Albert with m in {2,3,4},
graph i from seed 1000+i -- all in ONE persistent launch per step.  A step is
one vc_solve over the whole batch with the graphs already resident in HBM.
About half of the Erdos-Renyi graphs of this distribution are out of reach of
the reference's rule set (hours to years of search), so every graph carries a
tree-node budget; a graph that exhausts it is reported unsolved, never
dropped, and `value` counts exactly solved graphs only.

Other workloads (--workload): single_er (G(200,0.05) seed 0, plain MVC),
single_reg3 (3-regular n=120, 5% rewired), single_ba (BA n=2000 m=3, PVC at
k = opt-1 and opt).

Prints ONE JSON line (see the keys below).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "parallel-vertex-cover-algorithms-on-gpus_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

import graphgen  # noqa: E402
import vcgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0
KERNEL_SOURCES = [os.path.join(PKG, "csrc", f) for f in ("vc_kernels.hip", "vc_device.hpp")]
BATCH_PROFILE = os.path.join(ROOT, "profiles", "solve_kernel_batch4096.md")


def _code_only(text: str) -> str:
    """C/C++ source without comments and with whitespace runs collapsed, so a
    comment edit leaves the hash (and the compiled kernel) unchanged."""
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c in "\"'":  # string or character literal, kept whole
            j = i + 1
            while j < n and text[j] != c:
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
            out.append(" ")
        else:
            out.append(c)
            i += 1
    return " ".join("".join(out).split())


def pct(a, q):
    """q-th percentile of a sequence, None when it is empty."""
    return float(np.percentile(a, q)) if len(a) else None


def kernel_source_sha() -> str:
    """Hash of the kernel sources (code only) a committed profile was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, encoding="utf-8") as fh:
            h.update(_code_only(fh.read()).encode())
    return h.hexdigest()[:16]


def committed_traffic(path: str = BATCH_PROFILE):
    """HBM bytes per launch of the solve kernel on the batched workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary; None when there
    is none or it was measured on other kernel sources."""
    import re
    try:
        with open(path) as fh:
            text = fh.read()
    except OSError:
        return None, None
    kv = dict(re.findall(r"^- (\w+): `([^`]*)`", text, re.M))
    if kv.get("kernel_source_sha256") != kernel_source_sha() or "traffic_bytes_per_launch" not in kv:
        return None, "stale"
    return float(kv["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)


def _rocpd(db, sql):
    import sqlite3
    con = sqlite3.connect(db)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def _dispatches(db):
    """(kernel, dispatch id, duration ns) in start order."""
    rows = _rocpd(db, "select s.kernel_name, d.id, d.end - d.start from rocpd_kernel_dispatch d "
                      "join rocpd_info_kernel_symbol s on s.id = d.kernel_id order by d.start")
    return [(k.split("(")[0], i, dur) for k, i, dur in rows]


def _pmc(db):
    """{(dispatch id, counter): value} summed over the counter's instances."""
    rows = _rocpd(db, "select d.id, p.name, e.value from rocpd_pmc_event e join rocpd_info_pmc p on p.id = e.pmc_id "
                      "join rocpd_kernel_dispatch d on d.event_id = e.event_id")
    out = {}
    for i, name, v in rows:
        out[(i, name)] = out.get((i, name), 0.0) + v
    return out


def rocpd_summary(dbs, long_ms: float = 100.0):
    """Markdown lines for rocprofv3 result databases: per-kernel dispatch
    counts and mean durations, every dispatch longer than long_ms, and the
    counters of those dispatches."""
    lines = []
    for db in dbs:
        lines.append(f"### `{os.path.basename(db)}`")
        ds = _dispatches(db)
        pmc = _pmc(db)
        per = {}
        for k, i, dur in ds:
            per.setdefault(k, []).append((i, dur))
        for k, v in per.items():
            durs = [d for _, d in v]
            lines.append(f"- {k}: {len(v)} dispatches, mean {sum(durs) / len(durs) / 1e6:.3f} ms")
            for i, d in v:
                if d / 1e6 >= long_ms or "calib" in k:
                    ctr = ", ".join(f"{n} = {val:.6g}" for (j, n), val in sorted(pmc.items()) if j == i)
                    lines.append(f"  - dispatch {i}: {d / 1e6:.3f} ms" + (f"; {ctr}" if ctr else ""))
    return lines


def traffic_from(fetch_db, write_db, cal_fetch_db, cal_write_db, cal_bytes: int, kernel: str = "solve_kernel"):
    """HBM bytes of the last long dispatch of `kernel` (the timed launch):
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by the factor
    that makes the same counter read cal_bytes for the calibration copy
    kernel (which reads and writes cal_bytes per dispatch)."""
    def last(db, counter):
        pmc = _pmc(db)
        ds = [(i, dur) for k, i, dur in _dispatches(db) if kernel in k]
        i, dur = max(ds, key=lambda x: x[0]) if ds else (None, 0)
        return pmc.get((i, counter)), dur
    def calib(db, counter):
        pmc = _pmc(db)
        vals = [pmc.get((i, counter)) for k, i, _ in _dispatches(db) if "calib_copy" in k]
        vals = [v for v in vals if v]
        return cal_bytes / (sum(vals) / len(vals) * 1024.0)
    f_kib, f_dur = last(fetch_db, "FETCH_SIZE")
    w_kib, w_dur = last(write_db, "WRITE_SIZE")
    ff, wf = calib(cal_fetch_db, "FETCH_SIZE"), calib(cal_write_db, "WRITE_SIZE")
    total = f_kib * 1024.0 * ff + w_kib * 1024.0 * wf
    return {"fetch_kib": f_kib, "write_kib": w_kib, "fetch_factor": ff, "write_factor": wf,
            "fetch_dispatch_ms": f_dur / 1e6, "write_dispatch_ms": w_dur / 1e6, "traffic_bytes": total}


def b_node(n: int, m: int) -> int:
    """Algorithmic HBM bytes of one tree node: node state in (4n+4), child
    state out (4n+4), one CSR sweep (4(n+1) + 8m)."""
    return 12 * n + 8 * m + 12


def build_batch(count: int, base_seed: int):
    graphs, meta = [], []
    for i in range(count):
        n, e, desc = graphgen.batch_graph(i, base_seed)
        graphs.append(vcgpu.csr_from_edges(n, e))
        meta.append((n, len(e), desc))
    return graphs, meta


def cpu_baseline_batch(graphs_edges, budget_s: float, threads: int, max_nodes: int):
    """The oracle (restated reference CPU solver) over a bounded sample."""
    import oracle as O
    og = [O.from_edges(n, e) for n, e in graphs_edges]
    # grow the sample until about budget_s of CPU work has been done
    done_graphs, total_secs, solved = 0, 0.0, 0
    chunk = max(threads * 2, 8)
    i = 0
    while i < len(og) and total_secs < budget_s:
        part = og[i:i + chunk]
        secs, res, nodes, done = O.solve_batch_mt(part, threads, max_nodes)
        total_secs += secs
        done_graphs += len(part)
        solved += int(done.sum())
        i += chunk
    return {"value": solved / total_secs if total_secs > 0 else None, "unit": "graphs/s", "cores": threads,
            "kind": "port", "sample": f"first {done_graphs} graphs of the batch, oracle RemoveMax U + Sequential, "
                                      f"{threads} threads, per-graph cap {max_nodes} CPU tree nodes; {solved} solved "
                                      f"in {total_secs:.1f} s"}


def cpu_baseline_single(n, e, budget_s: float):
    import oracle as O
    g = O.from_edges(n, e)
    U = O.remove_max(g)
    t0 = time.perf_counter()
    # node budget sized from a short probe so the sample takes ~budget_s
    r, st, _ = O.sequential(g, U, max_nodes=20000)
    probe = time.perf_counter() - t0
    rate = st["nodes"] / max(probe, 1e-6)
    cap = int(rate * budget_s)
    t0 = time.perf_counter()
    r, st, _ = O.sequential(g, U, max_nodes=cap)
    secs = time.perf_counter() - t0
    return {"value": st["nodes"] / secs, "unit": "tree-nodes/s", "cores": 1, "kind": "port",
            "sample": f"oracle Sequential on the same graph, first {st['nodes']} tree nodes "
                      f"({'complete' if st['completed'] else 'capped'}) in {secs:.1f} s on 1 core"}


def served(ctx, args, out):
    """Served mixed stream: Poisson arrivals of MVC / PVC requests through
    two continuous-batching servers (small and large graph classes) that
    share the GPU.

    Small class (95%): graphs of the batch distribution, 50/50 MVC or PVC at
    k = opt +- 1, with opt from an offline batch solve of the same graphs (the
    "alone" answers); graphs whose opt that precompute cannot establish within
    its budget are not drawn, since their PVC requests need opt.

    Large class (5%): Barabasi-Albert n=20000 m=4.  Their exact MVC is far
    out of reach of this rule set (U ~ 10,800 and a search tree thousands of
    levels deep), so opt is unknown: they are served with a per-request node
    budget, 50/50 as MVC (returns an incumbent, counted unsolved when the
    budget ends the search) and PVC at k = U, whose answer is "yes" (the
    approximation's cover has size U); a "yes" is only counted when its cover
    checks.  `value` counts exactly answered requests of both classes."""
    import threading
    rng = np.random.Generator(np.random.PCG64(args.base_seed + 77))
    n_req = args.requests
    n_large = max(1, n_req // 20)
    # candidate pool and offline opt (small class)
    small_cands = [graphgen.batch_graph(i, args.base_seed)[:2] for i in range(min(3 * n_req, 4096))]
    large_cands = [(20000, graphgen.barabasi_albert(20000, 4, 5000 + j)) for j in range(min(n_large, 8))]
    pool = [vcgpu.csr_from_edges(n, e) for n, e in small_cands + large_cands]
    large_edges = {len(small_cands) + j: e for j, (_, e) in enumerate(large_cands)}
    ubs = [vcgpu.upper_bound(g, seed=args.base_seed) for g in pool]
    t0 = time.perf_counter()
    res_small, _, _ = ctx.solve(ctx.upload(pool[:len(small_cands)]),
                                [vcgpu.Job(vcgpu.VC_MVC, 0, u, 1 << 18) for u, _ in ubs[:len(small_cands)]],
                                vcgpu.Params(time_limit_ms=args.time_limit_ms), want_covers=False,
                                raise_on_error=False)
    precompute_s = time.perf_counter() - t0
    opts = [r.minimum if r.solved else None for r in res_small]
    small_ok = [i for i in range(len(small_cands)) if opts[i] is not None]
    large_ids = list(range(len(small_cands), len(pool)))
    # the stream: (graph, job, expected answer or None, class)
    reqs = []
    for r in range(n_req):
        big = rng.random() < 0.05
        mvc = rng.random() < 0.5
        if big:
            gi = int(rng.choice(large_ids))
            U = ubs[gi][0]
            if mvc:
                reqs.append((gi, vcgpu.Job(vcgpu.VC_MVC, 0, U, args.large_budget), None, "large"))
            else:
                reqs.append((gi, vcgpu.Job(vcgpu.VC_PVC, U, U, args.large_budget), ("pvc", 1), "large"))
            continue
        gi = int(small_ok[int(rng.integers(len(small_ok)))])
        opt = opts[gi]
        if mvc:
            reqs.append((gi, vcgpu.Job(vcgpu.VC_MVC, 0, ubs[gi][0], 0), ("mvc", opt), "small"))
        else:
            k = max(0, opt + (1 if rng.random() < 0.5 else -1))
            reqs.append((gi, vcgpu.Job(vcgpu.VC_PVC, k, ubs[gi][0], 0), ("pvc", int(k >= opt)), "small"))
    gaps = rng.exponential(1.0 / args.rate, size=n_req)
    # two servers: the small class takes most of each CU, the large class the rest
    # The large server's stacks hold a full path of any large request
    # (admission rejects requests needing more).
    large_depth = max(min(pool[gi].n, ubs[gi][0] + 1) + 1 for gi in large_ids)
    # The large server starts first: its waves need more registers, and a
    # persistent kernel launched second only gets the CUs' leftovers.
    large_srv = vcgpu.Server(ctx, max_vertices=20000, slots=64, arena_bytes=256 << 20, block_dim=64,
                             num_blocks=args.large_blocks, stack_depth=large_depth, time_limit_ms=args.time_limit_ms)
    small_srv = vcgpu.Server(ctx, max_vertices=512, slots=4096, arena_bytes=256 << 20,
                             num_blocks=args.small_blocks, time_limit_ms=args.time_limit_ms)
    arrivals = [0] * n_req
    tickets = [None] * n_req
    results = [None] * n_req

    def collect():
        for r in range(n_req):
            while tickets[r] is None:
                time.sleep(0.0002)
            srv, t = tickets[r]
            results[r] = srv.wait(t, args.time_limit_ms, want_cover=reqs[r][3] == "large")

    th = threading.Thread(target=collect)
    th.start()
    start = vcgpu.now_ns()
    due = start
    for r, (gi, job, want, cls) in enumerate(reqs):
        due += int(gaps[r] * 1e9)
        while True:
            left = due - vcgpu.now_ns()
            if left <= 0:
                break
            if left > 300_000:
                time.sleep((left - 200_000) / 1e9)
        srv = large_srv if cls == "large" else small_srv
        arrivals[r] = vcgpu.now_ns()
        tickets[r] = (srv, srv.submit(pool[gi], job))
    th.join()
    info_s = small_srv.stop()
    info_l = large_srv.stop()
    lat = {"small": [], "large": []}
    exact = {"small": 0, "large": 0}
    mism, budget_hit, last = 0, 0, start
    for r, (gi, job, want, cls) in enumerate(reqs):
        if results[r] is None:
            continue
        res, done_ns = results[r]
        lat[cls].append((done_ns - arrivals[r]) / 1e6)
        last = max(last, done_ns)
        if not res.solved or want is None:
            budget_hit += int(not res.solved)
            continue
        if want[0] == "mvc":
            ok = res.minimum == want[1]
        elif cls == "large":
            # k = U: "yes" only counts with a cover that checks
            e = large_edges[gi]
            cov = res.cover
            ok = (res.k_found == 1 and cov is not None and int(cov.sum()) <= job.k
                  and bool(np.all(cov[e[:, 0]] | cov[e[:, 1]])))
        else:
            ok = res.k_found == want[1]
        mism += int(not ok)
        exact[cls] += int(ok)
    wall = (last - start) / 1e9
    both = np.asarray(lat["small"] + lat["large"])
    solved = exact["small"] + exact["large"]
    out.update({
        "value": solved / wall, "unit": "graphs/s", "ms_per_step": wall * 1000.0,
        "config": {"workload": f"served stream, Poisson {args.rate:g} req/s, {n_req} requests, 95% batch-distribution "
                               f"graphs (MVC / PVC k=opt+-1) + 5% BA n=20000 m=4 (MVC / PVC k=U, node budget "
                               f"{args.large_budget})",
                   "requests": n_req, "exact": solved, "mismatches": mism, "unsolved_at_budget": budget_hit,
                   "small_requests": sum(1 for q in reqs if q[3] == "small"), "small_exact": exact["small"],
                   "large_requests": sum(1 for q in reqs if q[3] == "large"), "large_exact": exact["large"],
                   "p50_ms": pct(both, 50), "p99_ms": pct(both, 99), "mean_ms": float(both.mean()) if len(both) else None,
                   "p50_ms_small": pct(lat["small"], 50), "p99_ms_small": pct(lat["small"], 99),
                   "p50_ms_large": pct(lat["large"], 50), "p99_ms_large": pct(lat["large"], 99),
                   "precompute_s": precompute_s, "small_workers": info_s.workers, "large_workers": info_l.workers,
                   "large_stack_depth": large_depth,
                   "tree_nodes": info_s.tree_nodes + info_l.tree_nodes,
                   "idle_fraction_small": info_s.idle_fraction},
        "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                     "traffic": None},
    })


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "single_er", "single_reg3", "single_ba", "served"])
    ap.add_argument("--requests", type=int, default=2000, help="served: number of requests")
    ap.add_argument("--rate", type=float, default=1000.0, help="served: Poisson arrival rate (requests/s)")
    ap.add_argument("--small-blocks", type=int, default=0,
                    help="served: workgroups of the small-graph server (0 = the library default, 2 workers per CU)")
    ap.add_argument("--large-blocks", type=int, default=32,
                    help="served: one-wave workgroups of the large-graph server (each holds an exact-depth stack, "
                         "about 2 GB for BA n=20000)")
    ap.add_argument("--large-budget", type=int, default=1 << 12, help="served: node budget of a large request")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--base-seed", type=int, default=1000)
    ap.add_argument("--max-nodes", type=int, default=1 << 20, help="per-graph GPU tree-node budget (batch)")
    ap.add_argument("--plain-sample", type=int, default=256, help="graphs solved one per launch for the plain rate")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="batch: skip the autotuner and graph capture")
    ap.add_argument("--block-dim", type=int, default=0)
    ap.add_argument("--num-blocks", type=int, default=0)
    ap.add_argument("--time-limit-ms", type=int, default=120000)
    ap.add_argument("--list-size", type=int, default=0, help="worklist slots (-q), 0 = auto")
    ap.add_argument("--list-threshold", type=float, default=0.0, help="donation threshold fraction (-t), 0 = auto")
    ap.add_argument("--rocpd-summary", nargs="+", metavar="DB",
                    help="print a markdown summary of rocprofv3 result databases and exit (no GPU use)")
    ap.add_argument("--traffic-from", nargs=4, metavar=("FETCH_DB", "WRITE_DB", "CAL_FETCH_DB", "CAL_WRITE_DB"),
                    help="print the calibrated HBM traffic of the timed solve launch and exit (no GPU use)")
    ap.add_argument("--calib-bytes", type=int, default=1 << 28, help="bytes the calibration copy reads/writes")
    args = ap.parse_args()
    if args.rocpd_summary or args.traffic_from:
        if args.rocpd_summary:
            print("\n".join(rocpd_summary(args.rocpd_summary)))
        if args.traffic_from:
            t = traffic_from(*args.traffic_from, cal_bytes=args.calib_bytes)
            for k, v in t.items():
                print(f"- {k}: `{v:.6g}`")
            print(f"- traffic_bytes_per_launch: `{t['traffic_bytes']:.6g}`")
            print(f"- kernel_source_sha256: `{kernel_source_sha()}`")
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent graphs would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    ctx = vcgpu.Context(0)
    params = vcgpu.Params(version=vcgpu.VC_HYBRID, block_dim=args.block_dim, num_blocks=args.num_blocks,
                          time_limit_ms=args.time_limit_ms, list_size=args.list_size,
                          list_threshold=args.list_threshold)
    out = {"metric": "graphs/sec and tree-nodes/sec at exact MVC; p99 latency (batched vs plain)",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded generators)"}

    if args.workload == "served":
        served(ctx, args, out)
    elif args.workload == "batch":
        graphs, meta = build_batch(args.count, args.base_seed)
        ubs = [vcgpu.upper_bound(g, seed=args.base_seed) for g in graphs]
        jobs = [vcgpu.Job(vcgpu.VC_MVC, 0, u, args.max_nodes) for u, _ in ubs]
        t_up = time.perf_counter()
        gset = ctx.upload(graphs)
        upload_s = time.perf_counter() - t_up
        # launch shape: the autotuner on a 512-graph sample of the same
        # distribution (short node budget), then graph-captured dispatch
        tuned = None
        if not args.no_tune:
            t0 = time.perf_counter()
            ns = min(512, len(graphs))
            sample = ctx.upload(graphs[:ns])
            sjobs = [vcgpu.Job(vcgpu.VC_MVC, 0, u, 1 << 16) for u, _ in ubs[:ns]]
            best, _ = ctx.autotune(sample, sjobs, params)
            sample.free()
            params = vcgpu.Params(version=params.version, block_dim=best.block_dim, num_blocks=args.num_blocks,
                                  use_global_memory=best.use_global_memory, list_size=best.list_size,
                                  list_threshold=args.list_threshold, time_limit_ms=args.time_limit_ms,
                                  flags=vcgpu.VC_FLAG_GRAPH)
            tuned = {"block_dim": best.block_dim, "use_global_memory": best.use_global_memory,
                     "list_size": best.list_size, "autotune_s": time.perf_counter() - t0}
            print(f"bench: autotuned {tuned}", file=sys.stderr, flush=True)
        for i in range(args.warmup):
            ctx.solve(gset, jobs, params, want_covers=False, raise_on_error=False)
            print(f"bench: warmup {i} done", file=sys.stderr, flush=True)
        walls, kms, last = [], [], None
        for i in range(args.steps):
            t0 = time.perf_counter()
            res, info, st = ctx.solve(gset, jobs, params, want_covers=False, raise_on_error=False)
            walls.append(time.perf_counter() - t0)
            kms.append(info.kernel_ms)
            last = (res, info, st)
            print(f"bench: step {i} {walls[-1]:.3f} s, kernel {info.kernel_ms:.1f} ms, "
                  f"{sum(1 for r in res if r.solved)} solved", file=sys.stderr, flush=True)
        res, info, st = last
        solved = sum(1 for r in res if r.solved)
        nodes = sum(r.tree_nodes for r in res)
        wall = float(np.mean(walls))
        kernel_s = float(np.mean(kms)) / 1000.0
        alg_bytes = sum(r.tree_nodes * b_node(n, m) for r, (n, m, _) in zip(res, meta))
        traffic, traffic_src = (committed_traffic() if args.count == 4096 and args.max_nodes == 1 << 20
                                else (None, None))
        # plain: one graph per launch (default launch shape) on a sample,
        # every graph already resident in HBM before the clock starts
        ps = min(args.plain_sample, len(graphs))
        ones = [ctx.upload(graphs[i]) for i in range(ps)]
        plain_params = vcgpu.Params(version=vcgpu.VC_HYBRID, time_limit_ms=args.time_limit_ms)
        t0 = time.perf_counter()
        plain_solved, plain_lat = 0, []
        for i in range(ps):
            t1 = time.perf_counter()
            (r,), _, _ = ctx.solve(ones[i], jobs[i], plain_params, want_covers=False, raise_on_error=False)
            if r.solved:
                plain_lat.append((time.perf_counter() - t1) * 1000.0)
            plain_solved += int(r.solved)
        plain_wall = max(time.perf_counter() - t0, 1e-9)
        for one in ones:
            one.free()
        if ps:
            print(f"bench: plain {ps} graphs {plain_wall:.3f} s, {plain_solved} solved", file=sys.stderr, flush=True)
        out.update({
            "value": solved / wall, "unit": "graphs/s", "ms_per_step": wall * 1000.0,
            "config": {"workload": "batched MVC, 4096 graphs n~U[40,400], half ER p~U[0.02,0.15], half BA "
                                   "m in {2,3,4}, seed 1000+i, one persistent launch",
                       "graphs": len(graphs), "solved": solved, "unsolved_at_budget": len(graphs) - solved,
                       "per_graph_node_budget": args.max_nodes, "tree_nodes_per_step": nodes,
                       "tree_nodes_per_s": nodes / wall, "kernel_ms": kernel_s * 1000.0,
                       "workers": info.workers, "block_dim": info.block_dim, "num_blocks": info.num_blocks,
                       "blocks_per_cu": info.blocks_per_cu, "status": st,
                       # resident waves (one per worker; the persistent grid never exceeds
                       # occupancy) per SIMD, and as a fraction of the 8 wave slots of a SIMD
                       "waves_per_simd": info.workers / (info.cu_count * 4) if info.cu_count else None,
                       "wave_occupancy": info.workers / (info.cu_count * 4 * 8) if info.cu_count else None,
                       "idle_fraction": info.idle_fraction, "donations": info.donations, "steals": info.steals,
                       "tuned": tuned, "graph_captured": bool(params.flags & vcgpu.VC_FLAG_GRAPH),
                       "plain_graphs_per_s": plain_solved / plain_wall, "plain_sample": ps,
                       "plain_solved": plain_solved,
                       "upload_ms": upload_s * 1000.0,
                       "graphs_per_s_with_upload": solved / (wall + upload_s),
                       "batched_over_plain": ((solved / wall) / (plain_solved / plain_wall)
                                              if ps and plain_solved else None),
                       # latency of the solved graphs: batched = when a graph's search ended inside
                       # the one launch (all 4096 submitted together, measured from its first
                       # worker's start); plain = the wall time of one graph's own launch
                       "batched_p50_ms": pct([r.done_us / 1000.0 for r in res if r.solved], 50),
                       "batched_p99_ms": pct([r.done_us / 1000.0 for r in res if r.solved], 99),
                       "plain_p50_ms": pct(plain_lat, 50), "plain_p99_ms": pct(plain_lat, 99)},
            "roofline": {"bound": "hbm", "achieved": alg_bytes / kernel_s / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": alg_bytes / kernel_s / 1e9 / HBM_PEAK_GBS,
                         "traffic": traffic, "traffic_source": traffic_src,
                         "measured_hbm_gbs": traffic / kernel_s / 1e9 if traffic else None,
                         "algorithmic_bytes": alg_bytes, "kernel_ms": kernel_s * 1000.0},
        })
        if not args.no_cpu:
            th = args.cpu_threads or min(16, os.cpu_count() or 1)
            sample = [graphgen.batch_graph(i, args.base_seed)[:2] for i in range(min(len(graphs), 512))]
            out["cpu_baseline"] = cpu_baseline_batch(sample, args.cpu_seconds, th, args.max_nodes)
    else:
        n, e = graphgen.single(args.workload)
        g = vcgpu.csr_from_edges(n, e)
        U, _ = vcgpu.upper_bound(g, seed=0)
        gset = ctx.upload(g)
        if args.workload == "single_ba":
            # opt for k = opt - 1 / opt from an exact MVC solve (not timed)
            (r0,), _, _ = ctx.solve(gset, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0), params, want_covers=False)
            if not r0.solved:
                raise RuntimeError("single_ba: the MVC solve that fixes k did not finish")
            opt = r0.minimum
            jobs_list = [[vcgpu.Job(vcgpu.VC_PVC, opt - 1, U, 0)], [vcgpu.Job(vcgpu.VC_PVC, opt, U, 0)]]
        else:
            jobs_list = [[vcgpu.Job(vcgpu.VC_MVC, 0, U, 0)]]
        for _ in range(args.warmup):
            for jobs in jobs_list:
                ctx.solve(gset, jobs, params, want_covers=False)
        walls, kms, nodes, answers = [], [], 0, []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            step_nodes = 0
            step_k = 0.0
            for jobs in jobs_list:
                (r,), info, st = ctx.solve(gset, jobs, params, want_covers=False)
                step_nodes += r.tree_nodes
                step_k += info.kernel_ms
                answers.append(r.k_found if jobs[0].instance == vcgpu.VC_PVC else r.minimum)
            walls.append(time.perf_counter() - t0)
            kms.append(step_k)
            nodes = step_nodes
        wall = float(np.mean(walls))
        kernel_s = float(np.mean(kms)) / 1000.0
        out.update({
            "value": len(jobs_list) / wall, "unit": "graphs/s", "ms_per_step": wall * 1000.0,
            "config": {"workload": args.workload, "n": n, "m": len(e), "U": U, "answers": answers[-len(jobs_list):],
                       "tree_nodes_per_step": nodes, "tree_nodes_per_s": nodes / wall,
                       "kernel_ms": kernel_s * 1000.0, "workers": info.workers,
                       "idle_fraction": info.idle_fraction, "donations": info.donations, "steals": info.steals},
            "roofline": {"bound": "hbm", "achieved": nodes * b_node(n, len(e)) / kernel_s / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": nodes * b_node(n, len(e)) / kernel_s / 1e9 / HBM_PEAK_GBS, "traffic": None},
        })
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline_single(n, e, args.cpu_seconds)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
