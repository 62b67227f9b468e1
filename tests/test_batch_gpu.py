"""Batched and served solves: every value is the instance's exact max flow, identical alone
and in the mix."""
import time

import numpy as np
import pytest

from conftest import DEGENERATE, degenerate_graph, generated_graph, random_small_graph

pytestmark = pytest.mark.gpu


def mixed_instances(mf, golden, rng, k):
    names = ["rlg16_s3", "rlg32_s7", "genrmf8_s7", "genrmf12_s2", "rmat10_s5",
             "bip_500_700_s5", "bip_3000_2000_s9"]
    out = []
    for i in range(k):
        name = names[int(rng.integers(0, len(names)))]
        g, v = generated_graph(mf, golden, name)
        out.append((g, v))
    return out


def test_batch_matches_alone(mf, orc, ctx, golden):
    rng = np.random.default_rng(3)
    inst = mixed_instances(mf, golden, rng, 24)
    graphs = [g for g, _ in inst]
    values, res = ctx.batch_solve(graphs)
    assert values.tolist() == [v for _, v in inst]
    alone = [ctx.solve(g, mf.ALGO_VC).value for g in graphs[:6]]
    assert alone == values[:6].tolist()


def test_batch_bipartite_log_uniform(mf, orc, ctx):
    rng = np.random.default_rng(12)
    graphs = []
    for i in range(64):
        L, R = (int(round(np.exp(rng.uniform(np.log(64), np.log(2048))))) for _ in range(2))
        graphs.append(mf.gen_bipartite(L, R, 8.0, seed=1000 + i))
    expect = [orc.dinic_maxflow(g) for g in graphs]
    for mode in (0, 1):  # one workgroup per instance / packed graph on the whole grid
        for algo in (0, 1):  # TC / VC
            values, _ = ctx.batch_solve(graphs, batch_algo=algo, batch_mode=mode)
            assert values.tolist() == expect, (mode, algo)


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_mixed_generators_modes(mf, ctx, golden, mode):
    rng = np.random.default_rng(5 + mode)
    inst = mixed_instances(mf, golden, rng, 40)
    for algo in (0, 1):
        values, _ = ctx.batch_solve([g for g, _ in inst], batch_algo=algo, batch_mode=mode,
                                    tile_lanes=0)
        assert values.tolist() == [v for _, v in inst], (mode, algo)


@pytest.mark.parametrize("hops", [-1, 32])
def test_batch_forwarding(mf, ctx, golden, hops):
    rng = np.random.default_rng(9)
    inst = mixed_instances(mf, golden, rng, 24)
    for algo in (0, 1):
        values, _ = ctx.batch_solve([g for g, _ in inst], batch_algo=algo, forward_hops=hops)
        assert values.tolist() == [v for _, v in inst], (algo, hops)


def test_batch_fixed_cadence_and_limits(mf, ctx, golden):
    g, v = generated_graph(mf, golden, "genrmf12_s2")
    values, _ = ctx.batch_solve([g, g, g], cycles_per_relabel=8)
    assert values.tolist() == [v] * 3
    with pytest.raises(mf.MFError) as ei:
        ctx.batch_solve([g], max_rounds=1, cycles_per_relabel=1)
    assert ei.value.name == "MF_ERR_ITERATION_LIMIT"


def test_batch_single_and_errors(mf, ctx, golden):
    g, v = generated_graph(mf, golden, "genrmf8_s7")
    values, _ = ctx.batch_solve([g])
    assert values.tolist() == [v]
    bad = g.with_terminals(2, 2)
    with pytest.raises(mf.MFError):
        ctx.batch_solve([g, bad])


def test_server_continuous_batching(mf, ctx, golden):
    rng = np.random.default_rng(21)
    inst = mixed_instances(mf, golden, rng, 40)
    srv = mf.Server(ctx, hbm_budget_bytes=64 << 20, max_inflight_launches=3)
    ids = []
    try:
        for i, (g, v) in enumerate(inst):
            ids.append((srv.submit(g), v))
            if i % 5 == 4:
                srv.progress()  # requests keep arriving while earlier ones run
        pending = dict(ids)
        seen = []  # the completion queue hands out every id once, as it completes
        t0 = time.time()
        while pending and time.time() - t0 < 120:
            srv.progress(wait=True)
            for rid in srv.completed():
                seen.append(rid)
                assert srv.result(rid)[0] == pending.pop(rid)
        assert not pending
        assert sorted(seen) == sorted(rid for rid, _ in ids) and srv.completed() == []
        st = srv.stats()
        assert st["completed"] == len(inst) and st["submitted"] == len(inst)
        assert st["arena_bytes"] == 0 and 0 < st["arena_peak_bytes"] <= 64 << 20
    finally:
        srv.close()


def test_server_degenerate_graphs(mf, orc, ctx):
    # graphs without arcs, with only loops and backward arcs, zero capacities or int32-limit
    # capacities, served from the arena by the persistent kernel: every value equal to Dinic's
    graphs = [degenerate_graph(mf, k) for k in sorted(DEGENERATE)]
    expect = [orc.dinic_maxflow(g) for g in graphs]
    srv = mf.Server(ctx, hbm_budget_bytes=16 << 20, max_inflight_launches=2)
    try:
        ids = {srv.submit(g): v for g, v in zip(graphs, expect)}
        t0 = time.time()
        while ids and time.time() - t0 < 60:
            srv.progress(wait=True)
            for rid in srv.completed():
                assert srv.result(rid)[0] == ids.pop(rid)
        assert not ids
    finally:
        srv.close()


def test_server_random_small_graphs(mf, orc, ctx):
    # 80 small random multigraphs (capacities up to the int32 limit, degenerate ones among them)
    # submitted in bursts while earlier ones run: every served value equal to Dinic's
    rng = np.random.default_rng(31337)
    graphs = [random_small_graph(mf, rng) for _ in range(80)]
    expect = [orc.dinic_maxflow(g) for g in graphs]
    srv = mf.Server(ctx, hbm_budget_bytes=32 << 20, max_inflight_launches=2)
    try:
        ids = {}
        for i, (g, v) in enumerate(zip(graphs, expect)):
            ids[srv.submit(g)] = v
            if i % 10 == 9:
                srv.progress()
        t0 = time.time()
        while ids and time.time() - t0 < 60:
            srv.progress(wait=True)
            for rid in srv.completed():
                assert srv.result(rid)[0] == ids.pop(rid)
        assert not ids
    finally:
        srv.close()


def test_server_budget(mf, ctx, golden):
    # a 400 KB budget: a 200 KB arena (1,307 vertices, 13,070 slots) takes Genrmf 8x8 (513
    # vertices, 4,480 slots); Genrmf 16x16 fits neither the arena nor, alone (1.4 MB), the other
    # half of the budget, so it can never run
    g, v = generated_graph(mf, golden, "genrmf16_s1")
    budget = 400 << 10
    assert mf.footprint_bytes(g.n, g.m) > budget // 2
    srv = mf.Server(ctx, hbm_budget_bytes=budget, max_inflight_launches=1)
    try:
        with pytest.raises(mf.MFError) as ei:
            srv.submit(g)
        assert ei.value.name == "MF_ERR_BUDGET"
        h, hv = generated_graph(mf, golden, "genrmf8_s7")
        rid = srv.submit(h)
        srv.progress(wait=True)
        while srv.result(rid) is None:
            srv.progress(wait=True)
        assert srv.result(rid)[0] == hv
        assert srv.stats()["rejected"] == 1
    finally:
        srv.close()


def test_batch_autotune(mf, orc, ctx):
    rng = np.random.default_rng(44)
    graphs = []
    for i in range(48):
        L, R = (int(round(np.exp(rng.uniform(np.log(64), np.log(4096))))) for _ in range(2))
        graphs.append(mf.gen_bipartite(L, R, 8.0, seed=5000 + i))
    expect = [orc.dinic_maxflow(g) for g in graphs]
    best, ms, trials = ctx.batch_autotune(graphs, sample=16)
    assert trials >= 8 and ms > 0 and best["batch_algo"] in (0, 1)
    values, _ = ctx.batch_solve(graphs, **best)
    assert values.tolist() == expect


def test_server_large_requests_run_alone(mf, orc, ctx, golden):
    # requests above 16,384 vertices run alone on the whole grid, the others in batched
    # launches; every value is the instance's own
    big = [mf.gen_bipartite(20000, 15000, 8.0, seed=77), mf.gen_rmat(16, 16, seed=4)]
    expect_big = [orc.dinic_maxflow(g) for g in big]
    rng = np.random.default_rng(31)
    small = mixed_instances(mf, golden, rng, 12)
    srv = mf.Server(ctx, hbm_budget_bytes=1 << 30, max_inflight_launches=3)
    try:
        ids = []
        for i, (g, v) in enumerate(small):
            ids.append((srv.submit(g), v))
            if i in (3, 8):
                j = (i == 8)
                ids.append((srv.submit(big[j]), expect_big[j]))
        pending = dict(ids)
        t0 = time.time()
        while pending and time.time() - t0 < 120:
            srv.progress(wait=True)
            for rid in list(pending):
                r = srv.result(rid)
                if r is not None:
                    assert r[0] == pending.pop(rid)
        assert not pending
        st = srv.stats()
        assert st["completed"] == len(ids) and st["launches"] >= 3
    finally:
        srv.close()


def test_server_joins_running_batch(mf, orc, ctx):
    # requests submitted while earlier ones are still on the GPU are published into the same
    # running serving kernel: a second group starts before the first group has finished, and
    # one launch serves both
    g1 = [mf.gen_bipartite(8000, 8000, 8.0, seed=300 + i) for i in range(4)]
    g2 = [mf.gen_bipartite(8000, 8000, 8.0, seed=400 + i) for i in range(4)]
    expect = {id(g): orc.dinic_maxflow(g) for g in g1 + g2}
    srv = mf.Server(ctx, hbm_budget_bytes=1 << 30, max_inflight_launches=2)
    try:
        ids1 = [(srv.submit(g), g) for g in g1]
        t_end = time.time() + 0.003
        while time.time() < t_end:
            srv.progress()
        ids2 = [(srv.submit(g), g) for g in g2]
        pending = dict(ids1 + ids2)
        t0 = time.time()
        while pending and time.time() - t0 < 120:
            srv.progress(wait=True)
            for rid in list(pending):
                r = srv.result(rid)
                if r is not None:
                    assert r[0] == expect[id(pending.pop(rid))]
        assert not pending
        t1 = [srv.times(rid) for rid, _ in ids1]
        t2 = [srv.times(rid) for rid, _ in ids2]
        assert all(t["batched"] == 1 for t in t1 + t2)
        assert min(t["start_ms"] for t in t2) < max(t["finish_ms"] for t in t1)
        st = srv.stats()
        assert st["batched"] == 8 and st["serve_kernel_launches"] == 1
        # the library's latency figures are the request times' percentiles
        lat = np.array([t["finish_ms"] - t["submit_ms"] for t in t1 + t2])
        p50, p99, sps = srv.latency()
        assert abs(p50 - np.percentile(lat, 50)) < 1e-6 and abs(p99 - np.percentile(lat, 99)) < 1e-6
        assert sps > 0
    finally:
        srv.close()
