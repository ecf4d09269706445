"""Serving (continuous batching) on the GPU: each request's answer is the same
alone (its own one-graph launch) and in the mix (admitted into a running
persistent kernel with other requests), and equal to the oracle's.

Admission is forced to be continuous: fewer request slots than requests and a
small graph arena, so requests are admitted as earlier ones finish."""
import time

import numpy as np
import pytest

import graphgen
import oracle as O
import vcgpu

pytestmark = pytest.mark.gpu


def _requests():
    cases = []
    for s in range(12):
        n = 40 + 7 * s
        cases.append((n, graphgen.erdos_renyi(n, 0.08, 500 + s)))
    for s in range(8):
        n = 100 + 30 * s
        cases.append((n, graphgen.barabasi_albert(n, 2 + s % 3, 600 + s)))
    for s in range(4):
        cases.append((120, graphgen.regular3_rewired(120, 10 + s)))
    reqs = []
    for i, (n, e) in enumerate(cases):
        og = O.from_edges(n, e)
        opt, _, _ = O.mvc(og)
        g = vcgpu.csr_from_edges(n, e)
        U, ucov = vcgpu.upper_bound(g, seed=9)
        reqs.append((g, og, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0), opt, ucov))
        k = opt - 1 if i % 2 else opt + 1
        reqs.append((g, og, vcgpu.Job(vcgpu.VC_PVC, k, U, 0), opt, ucov))
    return reqs


def _check(res, og, job, opt, ucov):
    assert res.solved
    if job.instance == vcgpu.VC_MVC:
        assert res.minimum == opt
        cover = res.cover if res.cover is not None else ucov
        assert O.cover_is_valid(og, np.asarray(cover, np.uint8)) and int(np.sum(cover)) == opt
    else:
        assert res.k_found == int(job.k >= opt)
        if res.k_found:
            assert res.cover is not None
            assert O.cover_is_valid(og, np.asarray(res.cover, np.uint8)) and int(np.sum(res.cover)) <= job.k


def test_served_results_match_alone_and_oracle(ctx):
    reqs = _requests()
    # alone: one launch per request
    alone = []
    params = vcgpu.Params(time_limit_ms=60000)
    for g, og, job, opt, ucov in reqs:
        gs = ctx.upload(g)
        (r,), _, st = ctx.solve(gs, job, params)
        gs.free()
        _check(r, og, job, opt, ucov)
        alone.append(r)
    # in the mix: 8 slots and a 1 MiB arena for 48 requests, submitted in two waves
    srv = vcgpu.Server(ctx, max_vertices=512, slots=8, arena_bytes=1 << 20, time_limit_ms=60000)
    try:
        tickets = [srv.submit(g, job) for g, _, job, _, _ in reqs[:24]]
        time.sleep(0.05)
        tickets += [srv.submit(g, job) for g, _, job, _, _ in reqs[24:]]
        for t, (g, og, job, opt, ucov), a in zip(tickets, reqs, alone):
            out = srv.wait(t, 60000)
            assert out is not None, "request did not finish"
            r, done_ns = out
            _check(r, og, job, opt, ucov)
            if job.instance == vcgpu.VC_MVC:
                assert r.minimum == a.minimum
            else:
                assert r.k_found == a.k_found
    finally:
        info = srv.stop()
    assert info.tree_nodes > 0


def test_server_rejects_oversized_and_reports_pending(ctx):
    srv = vcgpu.Server(ctx, max_vertices=64, slots=4, time_limit_ms=30000)
    try:
        big = vcgpu.csr_from_edges(100, graphgen.erdos_renyi(100, 0.05, 1))
        with pytest.raises(vcgpu.VCError):
            srv.submit(big, vcgpu.Job(vcgpu.VC_MVC, 0, 100, 0))
        n, e = 50, graphgen.erdos_renyi(50, 0.1, 2)
        g = vcgpu.csr_from_edges(n, e)
        U, _ = vcgpu.upper_bound(g)
        t = srv.submit(g, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0))
        out = srv.wait(t, 30000)
        assert out is not None and out[0].solved
        assert out[0].minimum == O.mvc(O.from_edges(n, e))[0]
        assert srv.poll(t) is None  # a ticket is forgotten once its result was taken
    finally:
        srv.stop()


def test_server_admits_by_stack_footprint(ctx):
    """A request whose root-to-leaf path may need more stack entries than
    the server's stack_depth is refused at submit (its local stack could
    otherwise overflow into a full worklist); one that fits is served."""
    srv = vcgpu.Server(ctx, max_vertices=200, slots=4, stack_depth=40, time_limit_ms=30000)
    try:
        n, e = 120, graphgen.regular3_rewired(120, 3)
        g = vcgpu.csr_from_edges(n, e)
        U, _ = vcgpu.upper_bound(g)
        assert min(n, U) + 1 > 40
        with pytest.raises(vcgpu.VCError) as ei:
            srv.submit(g, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0))
        assert ei.value.status == vcgpu.VC_ENOTSUP
        n2, e2 = 60, graphgen.erdos_renyi(60, 0.05, 4)
        g2 = vcgpu.csr_from_edges(n2, e2)
        U2, _ = vcgpu.upper_bound(g2)
        assert min(n2, U2) + 1 <= 40
        out = srv.wait(srv.submit(g2, vcgpu.Job(vcgpu.VC_MVC, 0, U2, 0)), 30000)
        assert out is not None and out[0].solved and out[0].minimum == O.mvc(O.from_edges(n2, e2))[0]
    finally:
        srv.stop()


def test_hbm_server_beside_a_full_small_server(ctx):
    """Two persistent servers share the GPU: one whose node storage is in
    HBM (max_vertices 20000 does not fit LDS) and a small-graph server with
    the full occupancy grid (asked for explicitly: the default grid is
    smaller).  The larger-footprint server starts first (a persistent kernel
    launched second only gets the CUs left over).  Both answer exactly while
    both run, and stopping one does not wait for the other's kernel."""
    probe = vcgpu.Server(ctx, max_vertices=512, slots=64, time_limit_ms=60000)
    pi = probe.stop()
    full_blocks = pi.blocks_per_cu * pi.cu_count
    assert pi.num_blocks < full_blocks  # the default grid leaves CUs free
    big = []
    for s in range(2):
        n, e = 3000, graphgen.barabasi_albert(3000, 3, 900 + s)
        og = O.from_edges(n, e)
        opt, _, _ = O.mvc(og)
        g = vcgpu.csr_from_edges(n, e)
        U, _ = vcgpu.upper_bound(g, seed=9)
        big.append((g, og, U, opt))
    depth = max(min(g.n, U + 1) + 1 for g, _, U, _ in big)  # PVC at k = opt = U needs U + 2
    large = vcgpu.Server(ctx, max_vertices=20000, slots=8, block_dim=64, num_blocks=8, stack_depth=depth,
                         time_limit_ms=60000)
    small = vcgpu.Server(ctx, max_vertices=512, slots=64, num_blocks=full_blocks, time_limit_ms=60000)
    try:
        reqs = _requests()[:16]
        ts = [small.submit(g, job) for g, _, job, _, _ in reqs]
        tl = [(large.submit(g, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0)),
               large.submit(g, vcgpu.Job(vcgpu.VC_PVC, opt - 1, U, 0)),
               large.submit(g, vcgpu.Job(vcgpu.VC_PVC, opt, U, 0))) for g, _, U, opt in big]
        for t, (g, og, job, opt, ucov) in zip(ts, reqs):
            out = small.wait(t, 60000)
            assert out is not None
            _check(out[0], og, job, opt, ucov)
        for (tm, tno, tyes), (g, og, U, opt) in zip(tl, big):
            outs = [large.wait(t, 60000) for t in (tm, tno, tyes)]
            assert all(o is not None for o in outs), "large request did not finish"
            rm, rno, ryes = (o[0] for o in outs)
            assert rm.solved and rm.minimum == opt
            if rm.cover is not None:
                assert O.cover_is_valid(og, np.asarray(rm.cover, np.uint8)) and int(np.sum(rm.cover)) == opt
            assert rno.solved and rno.k_found == 0
            assert ryes.solved and ryes.k_found == 1 and ryes.cover is not None
            assert O.cover_is_valid(og, np.asarray(ryes.cover, np.uint8)) and int(np.sum(ryes.cover)) <= opt
    finally:
        t_stop = time.monotonic()
        info_l = large.stop()
        t_stop = time.monotonic() - t_stop
        small.stop()
    assert info_l.use_global_memory == 1
    assert t_stop < 10.0, t_stop  # the small server's kernel runs on (60 s limit)


def test_solve_beside_a_default_server_does_not_wait(ctx):
    """A batch solve issued while a server with the default grid runs gets
    CUs of its own and returns at once (a server that filled every CU would
    hold it until the server's time limit)."""
    n, e = 120, graphgen.regular3_rewired(120, 3)
    og = O.from_edges(n, e)
    opt, _, _ = O.mvc(og)
    g = vcgpu.csr_from_edges(n, e)
    U, _ = vcgpu.upper_bound(g, seed=1)
    gs = ctx.upload(g)
    srv = vcgpu.Server(ctx, max_vertices=512, slots=64, time_limit_ms=60000)
    try:
        t = time.monotonic()
        (r,), info, st = ctx.solve(gs, vcgpu.Job(vcgpu.VC_MVC, 0, U, 0), vcgpu.Params(time_limit_ms=30000),
                                  raise_on_error=False)
        t = time.monotonic() - t
    finally:
        srv.stop()
    assert r.solved and r.minimum == opt
    assert t < 10.0, t
