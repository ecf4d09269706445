"""The served stream's scoring (workloads.score) on hand-made results: every
request class and outcome lands in the count it belongs to, and every wrong
answer, cover that does not check, or incumbent above U is a mismatch.
No GPU: the results are constructed here."""
import numpy as np

import graphgen
import vcgpu
import workloads as W


def _stream(requests, graphs_edges, ubs):
    edges = [np.asarray(e, np.uint32).reshape(-1, 2) for e in graphs_edges]
    return W.Stream(graphs=[None] * len(edges), edges=edges, ubs=ubs, alone=[None] * len(edges),
                    requests=requests, gaps=np.zeros(len(requests)), precompute_s=0.0)


def _res(minimum, k_found=0, solved=True, cover=None):
    return vcgpu.Result(minimum=minimum, k_found=k_found, solved=solved, tree_nodes=1,
                        status=vcgpu.VC_OK if solved else vcgpu.VC_ETIMEOUT,
                        cover=None if cover is None else np.asarray(cover, bool))


def test_score_counts_every_outcome():
    # graph 0: path 0-1-2 (MVC 1: {1}); graph 1: triangle (MVC 2), U = 3
    path = [(0, 1), (1, 2)]
    tri = [(0, 1), (1, 2), (0, 2)]
    ubs = [(2, np.array([1, 0, 1], bool)), (3, np.array([1, 1, 1], bool))]
    R = W.Request
    reqs = [
        R(0, vcgpu.Job(vcgpu.VC_MVC, 0, 2, 0), "small", "mvc", 1),          # exact
        R(0, vcgpu.Job(vcgpu.VC_MVC, 0, 2, 0), "small", "mvc", 1),          # wrong size: mismatch
        R(0, vcgpu.Job(vcgpu.VC_PVC, 0, 2, 0), "small", "pvc", 0),          # k = opt-1: no, exact
        R(0, vcgpu.Job(vcgpu.VC_PVC, 1, 2, 0), "small", "pvc", 1),          # yes with a bad cover: mismatch
        R(1, vcgpu.Job(vcgpu.VC_MVC, 0, 3, 16), "small", "mvc_budget", None),  # budget ended, incumbent checks
        R(1, vcgpu.Job(vcgpu.VC_PVC, 3, 3, 16), "large", "large_pvc_at_U", 1),  # trivial yes
        R(1, vcgpu.Job(vcgpu.VC_MVC, 0, 3, 16), "large", "large_mvc", None),    # improved incumbent, unsolved
        R(1, vcgpu.Job(vcgpu.VC_MVC, 0, 3, 16), "large", "large_mvc", None),    # incumbent above U: mismatch
        R(0, vcgpu.Job(vcgpu.VC_MVC, 0, 2, 0), "small", "mvc", 1),          # no result: mismatch
    ]
    st = _stream(reqs, [path, tri], ubs)
    t0 = 1_000_000
    results = [
        (_res(1, cover=[0, 1, 0]), t0 + 1_000_000),
        (_res(2, cover=[1, 0, 1]), t0 + 2_000_000),
        (_res(0, k_found=0), t0 + 3_000_000),
        (_res(1, k_found=1, cover=[1, 0, 0]), t0 + 4_000_000),
        (_res(3, solved=False, cover=None), t0 + 5_000_000),
        (_res(3, k_found=1, cover=[1, 1, 1]), t0 + 6_000_000),
        (_res(2, solved=False, cover=[1, 1, 0]), t0 + 7_000_000),
        (_res(4, solved=False, cover=None), t0 + 8_000_000),
        None,
    ]
    arrivals = [t0] * len(reqs)
    out, wall, bad = W.score(st, results, arrivals, t0)
    small, large = out["small"], out["large"]
    assert small["n"] == 6 and large["n"] == 3
    assert small["exact"] == 2          # the MVC and the PVC no
    assert small["unsolved"] == 1       # the budget-ended MVC, checked against U's cover
    assert small["mismatches"] == 3     # wrong size, bad PVC cover, no result
    assert large["trivial_yes"] == 1 and large["exact"] == 0
    assert large["unsolved"] == 1 and large["improved"] == 1
    assert large["mismatches"] == 1     # incumbent above U
    assert len(bad) == 4
    assert abs(wall - 8.0 / 1000) < 1e-9  # last answer 8 ms after the start
    assert sorted(small["lat"]) == [1.0, 2.0, 3.0, 4.0, 5.0]


def test_cover_checks():
    e = np.asarray([(0, 1), (1, 2)], np.uint32)
    assert W.cover_checks(e, [0, 1, 0])
    assert not W.cover_checks(e, [1, 0, 0])
    assert W.cover_checks(e, [1, 1, 0], k=2) and not W.cover_checks(e, [1, 1, 1], k=2)
    assert not W.cover_checks(e, None)
    assert W.cover_checks(np.zeros((0, 2), np.uint32), [0, 0])


def test_stream_composition():
    # a pool of 24 batch graphs with made-up batch-solve results (every third
    # one unsolved), one large graph; no device is needed when the pool's
    # results are given
    pool = 24
    small = [graphgen.batch_graph(i, 1000)[:2] for i in range(pool)]
    graphs = [vcgpu.csr_from_edges(n, e) for n, e in small]
    ubs = [vcgpu.upper_bound(g, seed=1000) for g in graphs]
    alone = [_res(u - 1, solved=(i % 3 != 0)) for i, (u, _) in enumerate(ubs)]
    n_req, rate = 2000, 1000.0
    st = W.build_stream(None, n_req, rate, pool=pool, large_graphs=1, small_budget=1 << 16, large_budget=1 << 12,
                        pool_graphs=graphs, pool_ubs=ubs, pool_alone=alone)
    again = W.build_stream(None, n_req, rate, pool=pool, large_graphs=1, small_budget=1 << 16,
                           large_budget=1 << 12, pool_graphs=graphs, pool_ubs=ubs, pool_alone=alone)
    assert [(r.gi, r.kind, r.job.k) for r in st.requests] == [(r.gi, r.kind, r.job.k) for r in again.requests]
    assert len(st.requests) == n_req and len(st.gaps) == n_req
    assert abs(st.gaps.mean() * rate - 1.0) < 0.1          # Poisson arrivals at the stated rate
    large = [r for r in st.requests if r.cls == "large"]
    assert 60 <= len(large) <= 140                         # 5% of 2,000 (binomial, > 4 sigma)
    U_large = st.ubs[pool][0]
    assert st.graphs[pool].n == W.LARGE_N and len(st.edges[pool]) == W.LARGE_N * W.LARGE_M - W.LARGE_M ** 2
    for r in large:
        assert r.gi == pool and r.job.upper_bound == U_large and r.job.max_nodes == 1 << 12
        if r.kind == "large_mvc":
            assert r.job.instance == vcgpu.VC_MVC and r.expect is None
        else:
            assert r.kind == "large_pvc_at_U" and r.job.instance == vcgpu.VC_PVC
            assert r.job.k == U_large and r.expect == 1
    kinds = {}
    for r in st.requests:
        if r.cls != "small":
            continue
        kinds[r.kind] = kinds.get(r.kind, 0) + 1
        assert 0 <= r.gi < pool and r.job.max_nodes == 1 << 16 and r.job.upper_bound == ubs[r.gi][0]
        opt = alone[r.gi].minimum
        if not alone[r.gi].solved:
            assert r.kind == "mvc_budget" and r.job.instance == vcgpu.VC_MVC and r.expect is None
        elif r.kind == "mvc":
            assert r.job.instance == vcgpu.VC_MVC and r.expect == opt
        else:
            # PVC one below (answer no) or one above (answer yes) the known MVC
            assert r.kind == "pvc" and r.job.instance == vcgpu.VC_PVC
            assert r.job.k in (opt - 1, opt + 1) and r.expect == int(r.job.k >= opt)
    assert kinds.get("mvc", 0) > 0 and kinds.get("pvc", 0) > 0 and kinds.get("mvc_budget", 0) > 0
