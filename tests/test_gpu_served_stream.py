"""The served mixed stream at its stated workload: Poisson arrivals; 95% of
the requests on graphs of the batched distribution (MVC, or PVC at
k = opt +- 1), 5% on Barabasi-Albert n = 20,000 m = 4 graphs; two
continuous-batching servers share the GPU (workloads.py describes the
stream and its stand-ins for the large class, whose MVC is out of reach).

Checks: every small answer the stream finishes equals its alone answer (the
same graph in an offline batch launch) and, on the sample the CPU oracle
finishes, the oracle's; every large MVC incumbent is <= U with a cover that
checks; every large PVC "yes" comes with a valid cover of size <= k; no
request is lost."""
import numpy as np
import pytest

import oracle as O
import vcgpu
import workloads as W

pytestmark = pytest.mark.gpu


def test_served_stream_full_size(ctx):
    st = W.build_stream(ctx, n_req=400, rate=200.0, pool=1024, large_graphs=2, small_budget=1 << 20,
                        large_budget=1 << 12, time_limit_ms=120000, seed=91)
    assert sum(q.cls == "large" for q in st.requests) >= 10
    results, arrivals, start, info_s, info_l, depth = W.run_stream(ctx, st, time_limit_ms=120000)
    out, wall, bad = W.score(st, results, arrivals, start)
    print({c: {k: v for k, v in d.items() if k != "lat"} for c, d in out.items()}, f"wall {wall:.2f} s",
          f"small p50/p99 {np.percentile(out['small']['lat'], [50, 99])} ms",
          f"large p50/p99 {np.percentile(out['large']['lat'], [50, 99])} ms")
    assert not bad, bad[:5]
    assert all(r is not None for r in results)
    assert out["small"]["exact"] > 0.5 * out["small"]["n"]
    assert out["large"]["trivial_yes"] + out["large"]["unsolved"] + out["large"]["exact"] == out["large"]["n"]
    # the oracle on small graphs it finishes within 2^16 of its own nodes
    checked = 0
    for r, q in enumerate(st.requests):
        if q.cls != "small" or not results[r][0].solved or checked >= 40:
            continue
        n = st.graphs[q.gi].n
        opt, stats, _ = O.sequential(O.from_edges(n, st.edges[q.gi]), st.ubs[q.gi][0], max_nodes=1 << 16)
        if not stats["completed"]:
            continue
        res = results[r][0]
        if q.job.instance == vcgpu.VC_MVC:
            assert res.minimum == opt, (r, res.minimum, opt)
        else:
            assert res.k_found == int(q.job.k >= opt), (r, q.job.k, opt)
        checked += 1
    assert checked >= 20, checked
