"""vc_solve_noalloc (launch scratch in the library's fixed device array, no
allocation) against vc_solve on the same sets, through the C ABI: the same
answers, the same covers' validity and sizes, on two calls in a row, the
second on a smaller set, and again after a vc_solve has run between them.

The bar is the parity bar of vc_solve: bit-exact cover sizes and PVC
answers (both are exact searches; their node counts and the covers they
return may differ, so covers are checked for validity and size)."""
import numpy as np
import pytest

import graphgen
import oracle as O
import vcgpu

pytestmark = pytest.mark.gpu


def _cases(count, seed, nmax):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(8, nmax))
        if i % 2:
            e = graphgen.barabasi_albert(n, int(rng.integers(2, 4)), seed * 1000 + i)
        else:
            e = graphgen.erdos_renyi(n, float(rng.uniform(0.05, 0.25)), seed * 1000 + i)
        out.append((n, e))
    return out


_OPT = {}


def _opt(n, e):
    """Exact MVC from the oracle, computed once per graph."""
    key = (n, graphgen.edge_checksum(e))
    if key not in _OPT:
        _OPT[key] = O.mvc(O.from_edges(n, e))[0]
    return _OPT[key]


def _jobs(graphs, cases):
    jobs, want = [], []
    for g, (n, e) in zip(graphs, cases):
        U, _ = vcgpu.upper_bound(g, seed=3)
        opt = _opt(n, e)
        jobs.append(vcgpu.Job(vcgpu.VC_MVC, 0, U, 0))
        want.append(("mvc", opt))
        k = max(opt - 1, 0)
        jobs.append(vcgpu.Job(vcgpu.VC_PVC, k, U, 0))
        want.append(("pvc", int(k >= opt)))
    return jobs, want


def _check(name, cases2, res, want):
    for i, ((n, e), r, (kind, w)) in enumerate(zip(cases2, res, want)):
        assert r.solved, (name, i)
        if kind == "mvc":
            assert r.minimum == w, (name, i, r.minimum, w)
        else:
            assert r.k_found == w, (name, i, r.k_found, w)
        if r.cover is not None:
            og = O.from_edges(n, e)
            assert O.cover_is_valid(og, np.asarray(r.cover, np.uint8)), (name, i)
            assert int(np.sum(r.cover)) == r.minimum, (name, i)


def _run(ctx, cases, params):
    graphs = [vcgpu.csr_from_edges(n, e) for n, e in cases]
    jobs, want = _jobs(graphs, cases)
    gs = ctx.upload([g for g in graphs for _ in (0, 1)])
    cases2 = [c for c in cases for _ in (0, 1)]
    ref, ref_info, st = ctx.solve(gs, jobs, params)
    assert st == vcgpu.VC_OK
    got, info, st = ctx.solve(gs, jobs, params, noalloc=True)
    assert st == vcgpu.VC_OK
    _check("vc_solve", cases2, ref, want)
    _check("vc_solve_noalloc", cases2, got, want)
    for a, b in zip(ref, got):
        assert (a.minimum, a.k_found, a.solved) == (b.minimum, b.k_found, b.solved)
    assert info.tree_nodes == sum(r.tree_nodes for r in got)
    return info


def test_noalloc_matches_vc_solve_twice_second_smaller(ctx):
    params = vcgpu.Params(time_limit_ms=60000)
    first = _cases(48, 5, 120)
    second = _cases(9, 6, 60)
    i1 = _run(ctx, first, params)
    i2 = _run(ctx, second, params)
    print(f"noalloc workers {i1.workers} / {i2.workers}, depth {i1.stack_depth} / {i2.stack_depth}")
    # the same sets once more, hybrid then stack-only, without vc_solve in between
    for version in (vcgpu.VC_HYBRID, vcgpu.VC_STACK_ONLY):
        p = vcgpu.Params(version=version, start_depth=3, time_limit_ms=60000)
        for cases in (first, second):
            graphs = [vcgpu.csr_from_edges(n, e) for n, e in cases]
            jobs, want = _jobs(graphs, cases)
            gs = ctx.upload([g for g in graphs for _ in (0, 1)])
            got, info, st = ctx.solve(gs, jobs, p, noalloc=True)
            assert st == vcgpu.VC_OK
            _check(f"noalloc v{version}", [c for c in cases for _ in (0, 1)], got, want)


def test_noalloc_refuses_what_does_not_fit(ctx):
    """A graph whose one-workgroup exact-depth stacks exceed the fixed
    scratch is refused with VC_ENOMEM; graph dispatch with VC_ENOTSUP."""
    n = 20000
    g = vcgpu.csr_from_edges(n, graphgen.barabasi_albert(n, 4, 3))
    U, _ = vcgpu.upper_bound(g, seed=1)
    gs = ctx.upload(g)
    (r,), info, st = ctx.solve(gs, vcgpu.Job(vcgpu.VC_MVC, 0, U, 1000), vcgpu.Params(time_limit_ms=10000),
                               raise_on_error=False, noalloc=True)
    assert st == vcgpu.VC_ENOMEM
    small = ctx.upload(vcgpu.csr_from_edges(4, [(0, 1), (1, 2), (2, 3)]))
    (r,), info, st = ctx.solve(small, vcgpu.Job(vcgpu.VC_MVC, 0, 2, 0), vcgpu.Params(flags=vcgpu.VC_FLAG_GRAPH),
                               raise_on_error=False, noalloc=True)
    assert st == vcgpu.VC_ENOTSUP
    (r,), info, st = ctx.solve(small, vcgpu.Job(vcgpu.VC_MVC, 0, 2, 0), vcgpu.Params(), noalloc=True)
    assert st == vcgpu.VC_OK and r.solved and r.minimum == 2
