"""mf_solve_noalloc: the same results as mf_solve, with every device array in the library's
static scratch array (no device allocation).  Two calls in a row, the second on a smaller graph
(the scratch keeps the first call's contents), each compared with mf_solve and certified."""
import numpy as np
import pytest

from conftest import DEGENERATE, degenerate_graph, generated_graph

pytestmark = pytest.mark.gpu


def check_pair(mf, orc, ctx, g, value, algo):
    r0, f0 = ctx.solve(g, algo, flow=True)
    r1, f1 = ctx.solve_noalloc(g, algo, flow=True)
    assert r0.value == r1.value == value
    for flow in (f0, f1):
        cert, cut = orc.certify(g, flow, value)
        assert cert == "ok" and cut == value


@pytest.mark.parametrize("algo", [0, 1])
def test_noalloc_matches_solve_twice(mf, orc, ctx, golden, algo):
    big, vbig = generated_graph(mf, golden, "genrmf64_s1")  # 262,144 vertices: near the limit
    small, vsmall = generated_graph(mf, golden, "rmat12_s3")
    print(f"noalloc algo {algo}: genrmf64 then rmat12", flush=True)
    check_pair(mf, orc, ctx, big, vbig, algo)
    check_pair(mf, orc, ctx, small, vsmall, algo)


def test_noalloc_limits(mf, ctx, golden):
    g, v = generated_graph(mf, golden, "genrmf16_s1")
    with pytest.raises(mf.MFError) as ei:  # BCSR only
        mf._check(mf._lib.mf_solve_noalloc(ctx._p, mf.C.byref(g.csr()), 1, mf.LAYOUT_RCSR, None,
                                           None, None), "solve_noalloc")
    assert ei.value.name == "MF_ERR_INVALID_ARGUMENT"
    # above the 64 MiB scratch array: a clear error, nothing runs
    big = mf.gen_rmat(18, 16, seed=3)
    with pytest.raises(mf.MFError) as ei:
        ctx.solve_noalloc(big)
    assert ei.value.name == "MF_ERR_OVERFLOW"
    assert ctx.solve_noalloc(g).value == v  # still usable afterwards


@pytest.mark.parametrize("name", sorted(DEGENERATE))
def test_noalloc_degenerate_graphs(mf, orc, ctx, name):
    # graphs without arcs, with only loops and backward arcs, zero or int32-limit capacities,
    # after larger graphs left their contents in the scratch array
    g = degenerate_graph(mf, name)
    value = orc.dinic_maxflow(g)
    print(f"noalloc {name}", flush=True)
    for algo in (0, 1):
        check_pair(mf, orc, ctx, g, value, algo)


def test_noalloc_no_arcs_null_arrays(mf, ctx):
    # a graph without arcs whose arc arrays are NULL, as a C caller may pass it
    c = degenerate_graph(mf, "no_arcs").csr()
    c.dst = None
    c.cap = None
    for algo in (mf.ALGO_TC, mf.ALGO_VC):
        r = mf._Result()
        mf._check(mf._lib.mf_solve_noalloc(ctx._p, mf.C.byref(c), algo, mf.LAYOUT_BCSR, None,
                                           mf.C.byref(r), None), "mf_solve_noalloc")
        assert r.value == 0
