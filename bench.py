"""Benchmark of the gfx950 vertex-cover solver on one MI355X.

Default run: every configuration of the workload in one JSON line.

  value      batched MVC (the headline): 4096 graphs, n ~ U[40, 400], even i
             Erdos-Renyi with p ~ U[0.02, 0.15], odd i Barabasi-Albert with
             m in {2,3,4}, graph i from seed 1000+i -- all in ONE persistent
             launch per step (a step = one vc_solve over the batch, graphs
             resident in HBM).  About a third of these graphs are out of
             reach of the reference's rule set within the per-graph tree-node
             budget; they are reported unsolved, never dropped, and `value`
             counts exactly solved graphs only.  The plain form (one launch
             per graph) is timed on a sample beside it.
  singles    G(200, 0.05) seed 0 (plain MVC), 3-regular n=120 5% rewired
             (hard MVC), BA n=2000 m=3 seed 1 (PVC at k = opt-1 and opt):
             GPU median of five solves, CPU (the oracle, 1 core) median of
             five or, for G(200, 0.05), a bounded sample.
  served     Poisson arrivals of MVC / PVC requests on batch-distribution
             graphs (95%) and Barabasi-Albert n=20000 m=4 graphs (5%)
             through continuous-batching servers (workloads.py).
  roofline   bytes this design must move through HBM per launch (node-state
             copies, counted by the kernel) over the kernel time, beside the
             measured traffic and VALU issue of the committed profile.
  cpu_baseline  the oracle (a restatement of the reference's Sequential.cpp)
             on a stratified sample of the batch, one thread per core of this
             process's CPU share, with a speed check against the reference's
             own compiled CPU code.

--workload single_er | single_reg3 | single_ba | served runs one part alone.
--gpus N with N > 1 exits: this build runs on one GPU.
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
KERNEL_SOURCES = [os.path.join(PKG, "csrc", f) for f in ("vc_kernels.hip", "vc_device.hpp", "vc_bound.hpp")]
# vc_kernels.hip holds the device code first; the host launcher below this
# marker does not change the kernels a profile measured
HOST_MARKER = "host launcher and C ABI"
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


def kernel_source_sha(sources=None) -> str:
    """Hash of the device code (comments dropped, host launcher cut off) a
    committed profile was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sources or KERNEL_SOURCES:
        with open(f, encoding="utf-8") as fh:
            text = fh.read()
        cut = text.find(HOST_MARKER)
        h.update(_code_only(text if cut < 0 else text[:cut]).encode())
    return h.hexdigest()[:16]


def committed_profile(path: str = BATCH_PROFILE):
    """Machine-read figures of the committed rocprofv3 summary of the solve
    kernel on the batched workload (`- key: `value`` lines): HBM traffic per
    launch from FETCH_SIZE / WRITE_SIZE, VALU busy, ...  Returns (figures,
    source); figures is None when there is no summary or it was measured on
    other kernel sources (source "stale")."""
    import re
    try:
        with open(path) as fh:
            text = fh.read()
    except OSError:
        return None, None
    kv = dict(re.findall(r"^- (\w+): `([^`]*)`", text, re.M))
    if kv.get("kernel_source_sha256") != kernel_source_sha():
        return None, "stale"
    out = {}
    for k, v in kv.items():
        try:
            out[k] = float(v)
        except ValueError:
            out[k] = v
    return out, os.path.relpath(path, ROOT)


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
    """The reference-shaped per-node byte figure, kept as a labelled legacy
    field: node state in (4n+4), child state out (4n+4), one CSR sweep
    (4(n+1) + 8m), as the reference's global-memory kernels move it.  It
    charges every node a full HBM round trip, which this LDS-resident design
    does not make (its achieved rate exceeds the HBM peak)."""
    return 12 * n + 8 * m + 12


def cpu_share() -> dict:
    """The host CPU this process may use: model, logical CPUs, and the thread
    count of the baseline = the process's CPU affinity, capped by
    OMP_NUM_THREADS when it is set (a shared GPU host allots each GPU a share
    of its cores that way)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"cpu_model": model, "cpu_count": os.cpu_count(), "affinity": aff,
            "threads": min(aff, cap) if cap > 0 else aff}


def cpu_baseline_batch(count: int, base_seed: int, budget_s: float, threads: int, max_nodes: int):
    """The oracle (RemoveMax U + restated Sequential.cpp) on a stratified
    sample of the batch: graph indices in the order i * 1237 mod count
    (1237 is prime, so every stretch of the order spans the whole batch),
    in chunks over a `threads`-worker pool until about budget_s seconds of
    wall time, each graph with the GPU's per-graph node budget."""
    import oracle as O
    order = [(i * 1237) % count for i in range(count)]
    chunk = max(threads * 2, 8)
    total_secs, solved, done_graphs, nodes = 0.0, 0, 0, 0
    i = 0
    while i < len(order) and total_secs < budget_s:
        part = [O.from_edges(*graphgen.batch_graph(j, base_seed)[:2]) for j in order[i:i + chunk]]
        secs, res, nd, done = O.solve_batch_mt(part, threads, max_nodes)
        total_secs += secs
        done_graphs += len(part)
        solved += int(done.sum())
        nodes += int(nd.sum())
        i += chunk
    return {"value": solved / total_secs if total_secs > 0 else None, "unit": "graphs/s", "cores": threads,
            "kind": "port", "tree_nodes_per_s": nodes / total_secs if total_secs > 0 else None,
            "sample": f"{done_graphs} of the {count} batch graphs, stratified (indices i*1237 mod {count}), oracle "
                      f"RemoveMax U + Sequential on {threads} threads, per-graph cap {max_nodes} CPU tree nodes; "
                      f"{solved} solved, {nodes} tree nodes in {total_secs:.1f} s"}


def reference_speed_check(n: int = 150, p: float = 0.05, seed: int = 3):
    """The oracle against the reference's own compiled CPU code
    (oracle/_ref/ref_driver, built from its src/*.cpp) on one graph: the same
    answer and node count, and the ratio of their Sequential() times (1 core
    each).  None when the reference build is not present."""
    import tempfile

    import oracle as O
    if not O.ref_available():
        return None
    e = graphgen.erdos_renyi(n, p, seed)
    g = O.from_edges(n, e)
    U = O.remove_max(g)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.txt")
        graphgen.write_graph_file(path, n, e)
        ref = O.ref_run(path, minimum=U, timeout=600)
    t0 = time.perf_counter()
    r, st, _ = O.sequential(g, U)
    port_s = time.perf_counter() - t0
    return {"graph": f"G({n}, {p}) seed {seed}", "U": U, "mvc": r, "nodes": st["nodes"],
            "ref_mvc": ref["result"], "ref_nodes": ref["nodes"], "port_s": port_s, "ref_s": ref["seconds"],
            "port_over_ref": port_s / ref["seconds"] if ref["seconds"] > 0 else None,
            "agree": ref["result"] == r and ref["nodes"] == st["nodes"]}


def single_leg(ctx, name: str, reps: int = 5, cpu_budget_s: float = 10.0, params=None):
    """One single-graph workload: GPU median of `reps` warm solves (kernel and
    wall time), CPU oracle on 1 core (median of `reps` full solves when the
    oracle's node count is small, else a bounded sample giving tree-nodes/s
    and the full solve's time estimated from the oracle's node count)."""
    import oracle as O
    n, e = graphgen.single(name)
    g = vcgpu.csr_from_edges(n, e)
    U, _ = vcgpu.upper_bound(g, seed=0)
    gs = ctx.upload(g)
    params = params or vcgpu.Params(time_limit_ms=120000)
    if name == "single_ba":
        (r0,), _, _ = ctx.solve(gs, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0), params, want_covers=False)
        if not r0.solved:
            raise RuntimeError("single_ba: the MVC solve that fixes k did not finish")
        opt = r0.minimum
        jobs = [vcgpu.Job(vcgpu.VC_PVC, opt - 1, U, 0), vcgpu.Job(vcgpu.VC_PVC, opt, U, 0)]
    else:
        jobs = [vcgpu.Job(vcgpu.VC_MVC, 0, U, 0)]
    for j in jobs:  # warm
        ctx.solve(gs, j, params, want_covers=False)
    walls, kms, nodes, answers = [], [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        step_k, step_n, ans = 0.0, 0, []
        for j in jobs:
            (r,), info, st = ctx.solve(gs, j, params, want_covers=False)
            assert r.solved
            step_k += info.kernel_ms
            step_n += r.tree_nodes
            ans.append(r.k_found if j.instance == vcgpu.VC_PVC else r.minimum)
        walls.append((time.perf_counter() - t0) * 1000.0)
        kms.append(step_k)
        nodes.append(step_n)
        answers = ans
    gs.free()
    og = O.from_edges(n, e)
    Uo = O.remove_max(og)
    # CPU: the oracle's own MVC / PVC solve of the same question, 1 core
    def cpu_once(max_nodes=0):
        t0 = time.perf_counter()
        tot, done = 0, True
        if name == "single_ba":
            for j in jobs:
                _, st, _ = O.sequential_parameterized(og, Uo, j.k, max_nodes)
                tot += st["nodes"]
                done &= bool(st["completed"])
        else:
            _, st, _ = O.sequential(og, Uo, max_nodes)
            tot, done = st["nodes"], bool(st["completed"])
        return time.perf_counter() - t0, tot, done
    secs, cnodes, done = cpu_once(max_nodes=200000)
    out = {"n": n, "m": int(len(e)), "U": U, "answers": answers, "gpu_ms_median": float(np.median(walls)),
           "gpu_kernel_ms_median": float(np.median(kms)), "gpu_tree_nodes": int(np.median(nodes)),
           "gpu_tree_nodes_per_s": float(np.median(nodes) / (np.median(kms) / 1000.0)) if np.median(kms) else None}
    if done:
        ts = [cpu_once()[0] for _ in range(reps)]
        out.update({"cpu_ms_median": float(np.median(ts)) * 1000.0, "cpu_tree_nodes": cnodes, "cpu_cores": 1,
                    "cpu_sample": f"full solve, median of {reps}"})
    else:
        rate = cnodes / max(secs, 1e-9)
        cap = int(rate * cpu_budget_s)
        secs, cnodes, _ = cpu_once(max_nodes=cap)
        rate = cnodes / secs
        full = _oracle_nodes(name)
        out.update({"cpu_tree_nodes_per_s": rate, "cpu_cores": 1, "cpu_tree_nodes_full": full,
                    "cpu_ms_estimated": (full / rate * 1000.0) if full else None,
                    "cpu_sample": f"first {cnodes} of the oracle's tree nodes in {secs:.1f} s on 1 core; the full "
                                  f"solve's node count is the reference-checked golden vector"})
    cpu_ms = out.get("cpu_ms_median") or out.get("cpu_ms_estimated")
    out["gpu_over_cpu"] = (cpu_ms / out["gpu_ms_median"]) if cpu_ms and out["gpu_ms_median"] else None
    return out


def _oracle_nodes(name: str):
    """Tree nodes of the oracle's full solve of a single-graph workload, from
    the committed golden vectors (checked against the reference's own code)."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "oracle_vectors.json")) as f:
            for r in json.load(f):
                if r["kind"] == "single" and r["spec"].get("name") == name:
                    return int(r["oracle"]["nodes"])
    except OSError:
        pass
    return None


def served_leg(ctx, args, pool_graphs=None, pool_ubs=None, pool_alone=None):
    """Config 5: the served mixed stream (workloads.py)."""
    import workloads as W
    st = W.build_stream(ctx, n_req=args.requests, rate=args.rate, base_seed=args.base_seed,
                        pool=args.count, large_graphs=args.large_graphs, small_budget=args.served_budget,
                        large_budget=args.large_budget, time_limit_ms=args.time_limit_ms,
                        pool_graphs=pool_graphs, pool_ubs=pool_ubs, pool_alone=pool_alone)
    results, arrivals, start, info_s, info_l, depth = W.run_stream(
        ctx, st, small_blocks=args.small_blocks, large_blocks=args.large_blocks, time_limit_ms=args.time_limit_ms)
    out, wall, bad = W.score(st, results, arrivals, start)
    lat_all = out["small"]["lat"] + out["large"]["lat"]
    exact = out["small"]["exact"] + out["large"]["exact"]
    capacity = None
    if args.capacity_rate > 0:
        # the same kind of stream offered far faster than it can be served:
        # the answered rate is the servers' capacity, not the offered load
        pool = args.count
        cst = W.build_stream(ctx, n_req=args.requests, rate=args.capacity_rate, base_seed=args.base_seed, pool=pool,
                             large_graphs=args.large_graphs, small_budget=args.served_budget,
                             large_budget=args.large_budget, time_limit_ms=args.time_limit_ms, seed=78,
                             pool_graphs=st.graphs[:pool], pool_ubs=st.ubs[:pool], pool_alone=st.alone[:pool])
        cres, carr, cstart, _, _, _ = W.run_stream(ctx, cst, small_blocks=args.small_blocks,
                                                   large_blocks=args.large_blocks, time_limit_ms=args.time_limit_ms)
        cout, cwall, cbad = W.score(cst, cres, carr, cstart)
        cl = cout["small"]["lat"] + cout["large"]["lat"]
        capacity = {"offered_rate": args.capacity_rate, "requests": args.requests, "wall_s": cwall,
                    "requests_per_s": args.requests / cwall,
                    "graphs_per_s": (cout["small"]["exact"] + cout["large"]["exact"]) / cwall,
                    "small_requests_per_s": cout["small"]["n"] / cwall,
                    "p50_ms": pct(cl, 50), "p99_ms": pct(cl, 99), "mismatches": len(cbad)}
    return {
        "workload": f"served stream: Poisson {args.rate:g} req/s, {args.requests} requests; 95% batch-distribution "
                    f"graphs (MVC, or PVC at k = opt +- 1 with opt from the batch solve; budget "
                    f"{args.served_budget} nodes), 5% BA n=20000 m=4 (MVC incumbent or PVC at k = U, budget "
                    f"{args.large_budget} nodes)",
        "graphs_per_s": exact / wall, "exact": exact, "wall_s": wall, "requests": args.requests,
        "p50_ms": pct(lat_all, 50), "p99_ms": pct(lat_all, 99),
        "small": {k: v for k, v in out["small"].items() if k != "lat"},
        "large": {k: v for k, v in out["large"].items() if k != "lat"},
        "p50_ms_small": pct(out["small"]["lat"], 50), "p99_ms_small": pct(out["small"]["lat"], 99),
        "p50_ms_large": pct(out["large"]["lat"], 50), "p99_ms_large": pct(out["large"]["lat"], 99),
        "mismatches": len(bad), "precompute_s": st.precompute_s, "small_workers": info_s.workers,
        "large_workers": info_l.workers if info_l else 0, "large_stack_depth": depth,
        "tree_nodes": info_s.tree_nodes + (info_l.tree_nodes if info_l else 0),
        "idle_fraction_small": info_s.idle_fraction,
        "capacity": capacity,
    }


def batch_leg(ctx, args, params, out):
    graphs, meta = [], []
    for i in range(args.count):
        n, e, desc = graphgen.batch_graph(i, args.base_seed)
        graphs.append(vcgpu.csr_from_edges(n, e))
        meta.append((n, len(e), desc))
    ubs = [vcgpu.upper_bound(g, seed=args.base_seed) for g in graphs]
    jobs = [vcgpu.Job(vcgpu.VC_MVC, 0, u, args.max_nodes) for u, _ in ubs]
    t_up = time.perf_counter()
    gset = ctx.upload(graphs)
    upload_s = time.perf_counter() - t_up
    # launch shape: the autotuner on a 512-graph sample of the same
    # distribution with the step's own node budget (shorter budgets end the
    # sample's launches before the per-node rate that decides a full step
    # settles), then graph-captured dispatch
    tuned = None
    if not args.no_tune:
        t0 = time.perf_counter()
        ns = min(512, len(graphs))
        sample = ctx.upload(graphs[:ns])
        sjobs = [vcgpu.Job(vcgpu.VC_MVC, 0, u, args.max_nodes) for u, _ in ubs[:ns]]
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
    legacy_bytes = sum(r.tree_nodes * b_node(n, m) for r, (n, m, _) in zip(res, meta))
    prof, prof_src = (committed_profile() if args.count == 4096 and args.max_nodes == 1 << 20 else (None, None))
    traffic = prof.get("traffic_bytes_per_launch") if prof else None
    # plain: one graph per launch (default launch shape) on a sample, every
    # graph already resident in HBM before the clock starts
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
    gset.free()
    state_bytes = info.state_bytes
    achieved = state_bytes / kernel_s / 1e9
    out.update({
        "value": solved / wall, "unit": "graphs/s", "ms_per_step": wall * 1000.0,
        "config": {"workload": "batched MVC, 4096 graphs n~U[40,400], half ER p~U[0.02,0.15], half BA "
                               "m in {2,3,4}, seed 1000+i, one persistent launch",
                   "graphs": len(graphs), "solved": solved, "unsolved_at_budget": len(graphs) - solved,
                   "per_graph_node_budget": args.max_nodes, "tree_nodes_per_step": nodes,
                   "tree_nodes_per_s": nodes / wall, "kernel_ms": kernel_s * 1000.0,
                   "active_ms": info.active_ms,
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
        # HBM roof.  achieved = the bytes this design must move through HBM per
        # launch -- the node-state slot copies of every stack push / pop and
        # worklist steal / donation (8 bytes per vertex of the graph per copy,
        # counted by the kernel) -- over the kernel's duration.  traffic = the
        # committed profile's calibrated FETCH_SIZE + WRITE_SIZE of one launch;
        # valu_busy = its SQ_ACTIVE_INST_VALU x 4 / SIMDs / cycles, the roof that
        # binds this kernel.
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "model": "node-state slot copies (stack push/pop, worklist steal, 2x donation), "
                              "8 B per vertex per copy, counted in the kernel",
                     "algorithmic_bytes": state_bytes, "kernel_ms": kernel_s * 1000.0,
                     "measured_hbm_gbs": traffic / kernel_s / 1e9 if traffic else None,
                     "measured_frac": traffic / kernel_s / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "valu_busy": prof.get("valu_busy") if prof else None,
                     "binding": "VALU issue" if prof and prof.get("valu_busy") else None,
                     "profile": prof_src,
                     "b_node_legacy": {"achieved_gbs": legacy_bytes / kernel_s / 1e9,
                                       "frac": legacy_bytes / kernel_s / 1e9 / HBM_PEAK_GBS,
                                       "algorithmic_bytes": legacy_bytes,
                                       "note": "B_node = 12n+8m+12 per node (the reference's global-memory "
                                               "kernels' traffic): charges each node a full HBM round trip "
                                               "this LDS-resident design does not make"}},
    })
    return graphs, ubs, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    choices=["all", "batch", "single_er", "single_reg3", "single_ba", "served"])
    ap.add_argument("--requests", type=int, default=2000, help="served: number of requests")
    ap.add_argument("--rate", type=float, default=1000.0, help="served: Poisson arrival rate (requests/s)")
    ap.add_argument("--served-budget", type=int, default=1 << 16,
                    help="served: node budget of a small request (the serving policy's per-request cap)")
    ap.add_argument("--small-blocks", type=int, default=0,
                    help="served: workgroups of the small-graph server (0 = the library default, 2 workers per CU)")
    ap.add_argument("--large-blocks", type=int, default=32,
                    help="served: one-wave workgroups of the large-graph server (each holds an exact-depth stack, "
                         "about 2 GB for BA n=20000)")
    ap.add_argument("--large-budget", type=int, default=1 << 12,
                    help="served: node budget of a large request (a PVC request at k = U reaches its first cover "
                         "within 2^15 nodes shared by the server's workers, not within 2^12)")
    ap.add_argument("--large-graphs", type=int, default=4, help="served: distinct BA n=20000 graphs")
    ap.add_argument("--capacity-rate", type=float, default=50000.0,
                    help="served: offered rate of a second stream that saturates the servers (0 = skip)")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--base-seed", type=int, default=1000)
    ap.add_argument("--max-nodes", type=int, default=1 << 20, help="per-graph GPU tree-node budget (batch)")
    ap.add_argument("--plain-sample", type=int, default=256, help="graphs solved one per launch for the plain rate")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-singles", action="store_true")
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

    if args.workload in ("all", "batch"):
        graphs, ubs, res = batch_leg(ctx, args, params, out)
        if args.workload == "all" and not args.no_singles:
            out["singles"] = {}
            for name in ("single_er", "single_reg3", "single_ba"):
                out["singles"][name] = single_leg(ctx, name)
                print(f"bench: {name} {out['singles'][name]}", file=sys.stderr, flush=True)
        if args.workload == "all" and not args.no_served:
            out["served"] = served_leg(ctx, args, graphs, ubs, res)
            print(f"bench: served {out['served']}", file=sys.stderr, flush=True)
        if not args.no_cpu:
            share = cpu_share()
            th = args.cpu_threads or share["threads"]
            cb = cpu_baseline_batch(args.count, args.base_seed, args.cpu_seconds, th, args.max_nodes)
            cb.update({"cpu_model": share["cpu_model"], "cpu_count": share["cpu_count"],
                       "affinity": share["affinity"], "reference_speed_check": reference_speed_check()})
            out["cpu_baseline"] = cb
    elif args.workload == "served":
        sv = served_leg(ctx, args)
        out.update({"value": sv["graphs_per_s"], "unit": "graphs/s", "ms_per_step": sv["wall_s"] * 1000.0,
                    "config": sv, "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                                               "unit": "GB/s", "frac": None, "traffic": None}})
    else:
        sg = single_leg(ctx, args.workload, reps=max(args.steps, 1), params=params)
        out.update({"value": len(sg["answers"]) / (sg["gpu_ms_median"] / 1000.0), "unit": "graphs/s",
                    "ms_per_step": sg["gpu_ms_median"], "config": dict(workload=args.workload, **sg),
                    "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": None, "traffic": None}})
        if not args.no_cpu:
            out["cpu_baseline"] = {"value": sg.get("cpu_tree_nodes_per_s") or
                                   (sg["cpu_tree_nodes"] / (sg["cpu_ms_median"] / 1000.0)),
                                   "unit": "tree-nodes/s", "cores": 1, "kind": "port", "sample": sg["cpu_sample"]}
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
