"""The batched and serving configurations at their full sizes.

Batched: 4096 bipartite matching networks, |L| and |R| log-uniform in [256, 65536], Poisson(8)
arcs per left vertex, unit capacities (the same seeded mix the benchmark times), solved in one
batched call; every value must equal the CPU oracle's (Dinic) and a sample must equal
single-instance solves of the same instances.

Serving: a seeded Poisson schedule of requests mixing the RLG, Genrmf and bipartite generators
with 1e3..1e6 vertices, replayed in real time through the server; every served value must equal
the same request solved alone, and each alone solve is certified (capacity, conservation,
value, and a cut of equal capacity) -- a full-size property check that needs no second solver.
The bipartite formulation follows the paper's Table III (WBPR_HPEC2025_CameraReady.pdf p.6);
serving has no reference counterpart.
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from wbpr_amd.workloads import batched_mix, poisson_arrivals, serving_requests

pytestmark = pytest.mark.gpu

WORKERS = min(16, len(os.sched_getaffinity(0)))


@pytest.fixture(scope="module")
def mix4096(mf):
    graphs = batched_mix(mf, 4096, 2024)
    print(f"\nbatched mix: {len(graphs)} instances, {sum(g.n for g in graphs)} vertices, "
          f"{sum(g.m for g in graphs)} arcs", flush=True)
    return graphs


@pytest.fixture(scope="module")
def mix4096_values(orc, mix4096):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(orc.dinic_maxflow, mix4096))


@pytest.mark.parametrize("algo", [0, 1])  # batch_algo: TC, VC cycles inside each workgroup
def test_batched_4096_full_mix(mf, ctx, mix4096, mix4096_values, algo):
    print(f"batched 4096 instances, batch_algo {algo}", flush=True)
    values, res = ctx.batch_solve(mix4096, batch_algo=algo)
    bad = [i for i, (a, b) in enumerate(zip(values.tolist(), mix4096_values)) if a != b]
    assert not bad, f"instances {bad[:8]} differ from the CPU oracle"
    assert res.solve_ms > 0 and res.pushes > 0


def test_batched_4096_sample_alone(mf, ctx, mix4096, mix4096_values):
    # the largest instances and a seeded spread of the rest, solved alone on the single-instance
    # path (graph resident, whole grid)
    order = np.argsort([g.m for g in mix4096])
    rng = np.random.default_rng(7)
    pick = sorted(set(order[-4:].tolist()) | set(rng.choice(len(mix4096), 12, replace=False).tolist()))
    for i in pick:
        for algo in (mf.ALGO_TC, mf.ALGO_VC):
            dg = ctx.upload(mix4096[i], mf.LAYOUT_BCSR)
            r = dg.solve(algo)
            dg.close()
            assert r.value == mix4096_values[i], (i, algo)


def test_serving_poisson_schedule_full_sizes(mf, orc, ctx):
    reqs, kinds = serving_requests(mf, 24, 99)
    sizes = [g.n for g in reqs]
    assert max(sizes) > 2e5 and min(sizes) < 1e4  # the schedule spans the size range
    arrivals = poisson_arrivals(len(reqs), 20.0, 100)
    srv = mf.Server(ctx, hbm_budget_bytes=8 << 30, max_inflight_launches=4)
    served, ids, done_ids = {}, {}, {}
    try:
        t0 = time.perf_counter()
        nxt = 0
        while len(served) < len(reqs):
            now = time.perf_counter() - t0
            assert now < 100, f"serving stalled: {len(served)} of {len(reqs)} done"
            while nxt < len(reqs) and arrivals[nxt] <= now:
                ids[srv.submit(reqs[nxt])] = nxt
                nxt += 1
            srv.progress()
            for rid, i in list(ids.items()):
                r = srv.result(rid)
                if r is not None:
                    served[i] = r[0]
                    done_ids[rid] = i
                    del ids[rid]
            time.sleep(0.001)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["arena_bytes"] == 0
        assert st["batched"] + st["solo"] == len(reqs) and st["batched"] > 0
        # the batched requests share running launches of the serving kernel
        assert st["serve_kernel_launches"] <= st["batched"]
        times = {i: srv.times(rid) for rid, i in done_ids.items()}
        assert all(t["submit_ms"] <= t["admit_ms"] <= t["start_ms"] <= t["finish_ms"]
                   for t in times.values())
    finally:
        srv.close()
    for i, g in enumerate(reqs):
        print(f"request {i}: {kinds[i]} n={g.n} m={g.m}", flush=True)
        dg = ctx.upload(g, mf.LAYOUT_BCSR)
        r = dg.solve(mf.ALGO_VC)
        flow = dg.read_flow(return_excess=True)
        dg.close()
        assert served[i] == r.value, (i, kinds[i], g.n)
        cert, cut = orc.certify(g, flow, r.value)
        assert cert == "ok" and cut == r.value, (i, kinds[i], cert, cut)
