"""The batched configuration at its full size: 4096 graphs (n ~ U[40, 400],
half Erdos-Renyi p ~ U[0.02, 0.15], half Barabasi-Albert m in {2,3,4}, graph
i from seed 1000 + i) in one persistent launch, with a per-graph node budget
so the launch stays short.  Size-independent checks on every graph, the
oracle on a deterministic sample."""
import numpy as np
import pytest

import graphgen
import oracle as O
import vcgpu

pytestmark = pytest.mark.gpu


def test_batched_configuration_every_answer_checks(ctx):
    graphs = [graphgen.batch_graph(i, 1000)[:2] for i in range(4096)]
    csrs = [vcgpu.csr_from_edges(n, e) for n, e in graphs]
    ubs = [vcgpu.upper_bound(g, seed=1000) for g in csrs]
    jobs = [vcgpu.Job(vcgpu.VC_MVC, 0, u, 1 << 16) for u, _ in ubs]
    gs = ctx.upload(csrs)
    res, info, st = ctx.solve(gs, jobs, vcgpu.Params(time_limit_ms=120000), want_covers=True,
                              raise_on_error=False)
    assert st in (vcgpu.VC_OK, vcgpu.VC_ETIMEOUT)
    solved = [i for i, r in enumerate(res) if r.solved]
    assert len(solved) > 2000, len(solved)
    # each graph's end time lies inside the launch (first worker start to kernel end)
    done = [res[i].done_us for i in solved]
    assert max(done) > 0 and max(done) <= info.kernel_ms * 1000.0 * 1.02 + 1000.0, (max(done), info.kernel_ms)
    for i, ((n, e), r, (U, ucov)) in enumerate(zip(graphs, res, ubs)):
        e = np.asarray(e, np.int64).reshape(-1, 2)
        if r.solved:
            cover = r.cover if r.cover is not None else ucov
            cover = np.asarray(cover, bool)
            assert bool(np.all(cover[e[:, 0]] | cover[e[:, 1]])), i  # a vertex cover
            assert int(cover.sum()) == r.minimum <= U, i                # of the reported size
        else:
            assert r.status == vcgpu.VC_ETIMEOUT and r.minimum <= U, i  # budget ended it
    # the reference's Sequential from the same U, where it finishes within
    # 2^16 of its own tree nodes
    checked = 0
    for i in solved[::16]:
        n, e = graphs[i]
        opt, stats, _ = O.sequential(O.from_edges(n, e), ubs[i][0], max_nodes=1 << 16)
        if stats["completed"]:
            assert res[i].minimum == opt, (i, res[i].minimum, opt)
            checked += 1
    assert checked >= 40, checked
