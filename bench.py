#!/usr/bin/env python3
"""Max-flow benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--algo tc|vc]
                    [--layout bcsr|rcsr|bcsr2] [--no-cpu-baseline] [--no-variants]

Workloads
  genrmf64 (default)  Genrmf a=b=64, 262,144 vertices: a step is one complete max-flow solve of the
                      resident graph (each solve restarts from the capacities on the device)
  rmat22, rlg64       the other single-instance configurations
  batched             4096 bipartite matching networks (|L|, |R| log-uniform in [256, 65536],
                      Poisson(8) arcs per left vertex) solved in one batched call; a step is one
                      batched solve; compared with one-by-one single-instance solves
  serving             Poisson arrivals of mixed requests (RLG, Genrmf, bipartite; 1e3..1e6
                      vertices) through the continuous-batching server; p50/p99 latency
  genrmf128           Genrmf a=b=128 (2,097,152 vertices), the size of the paper's Genrmf row
  copy                1 GiB device-to-device copies: the achievable HBM rate, and the known-size
                      kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters

Every value is checked exactly: single instances against their committed values, batched and
served results against solves of the same instances one by one.  Prints ONE JSON line (the
driver's contract) with a `roofline` entry for the dominant kernel (algorithmic bytes per launch
over the mean launch time, from HIP events around every launch on the solver's stream) and a
`cpu_baseline` entry (the reference's own serial solver, or the CPU oracle where noted, on a
bounded sample of the same workload).
"""
import argparse
import ctypes
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "max-flow wall time + residual edges/s (exact value); batched solves/s vs one-by-one"
HBM_PEAK_GBS = 8000.0
PROFILE_EVERY = 16  # cycle launches per launch timed with dispatch-packet events

# single-instance workloads and their exact max-flow values (same numbers as
# tests/golden/values.json, each computed by the CPU oracles)
WORKLOADS = {
    "genrmf64": dict(
        desc="Genrmf a=64 b=64 c1=1 c2=10000 seed=1 (262,144 vertices, 1,290,240 arcs)",
        gen=("gen_genrmf", (64, 64, 1, 10000, 1)), value=19996127, algo="vc",
        variants=("tc", "vc")),
    "rmat22": dict(
        desc="R-MAT scale 22 edge factor 16 a=.57 b=.19 c=.19 caps [1,1000] seed=1 "
             "(4,194,304 vertices, 65,244,556 arcs after dropping self-loops and duplicates)",
        gen=("gen_rmat", (22, 16, 0.57, 0.19, 0.19, 1000, 1)), value=18886759, algo="vc"),
    "genrmf128": dict(
        desc="Genrmf a=128 b=128 c1=1 c2=10000 seed=1 (2,097,152 vertices, 10,403,840 arcs)",
        gen=("gen_genrmf", (128, 128, 1, 10000, 1)), value=81048332, algo="vc",
        variants=("tc", "vc")),
    "rlg64": dict(
        desc="Washington RLG 64x64 r=10000 seed=1 (4,098 vertices, 12,224 arcs)",
        gen=("gen_washington_rlg", (64, 64, 10000, 1)), value=466088, algo="tc"),
}
CPU_SAMPLE = dict(desc="Genrmf a=14 b=14 c1=1 c2=10000 seed=1 (2,744 vertices, 12,740 arcs)",
                  gen=("gen_genrmf", (14, 14, 1, 10000, 1)), value=919190)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def splitmix(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def cpu_baseline_reference(mf):
    """Reference serial solver (built from its sources) on the bounded sample, one core."""
    ref = os.path.join(ROOT, "oracle", "_ref", "maxflow-serial")
    if not os.path.exists(ref):
        return None
    g = getattr(mf, CPU_SAMPLE["gen"][0])(*CPU_SAMPLE["gen"][1])
    tmp = tempfile.mkdtemp()
    try:
        path = os.path.join(tmp, "sample.max")
        g.save_dimacs(path)
        cmd = [ref, "-v", "2", "-f", path, "-s", "0", "-t", "1"]
        if shutil.which("taskset"):
            cmd = ["taskset", "-c", "0"] + cmd
        t0 = time.perf_counter()
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        dt = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    got = None
    for line in out.stdout.splitlines():
        if line.startswith("Max flow:"):
            got = int(line.split()[-1])
    if got != CPU_SAMPLE["value"]:
        raise SystemExit(f"cpu baseline value {got} != {CPU_SAMPLE['value']}")
    return {"value": round(2 * g.m / dt, 1), "unit": "residual edges/s", "cores": 1,
            "kind": "reference",
            "sample": f"maxflow-serial (reference sources) on {CPU_SAMPLE['desc']}: "
                      f"{dt:.2f} s wall, pinned to core 0; value = residual slots / time"}


def kernel_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def roofline_of(results, workload):
    """Roofline of the dominant kernel over solves run with profile_kernels=1: algorithmic bytes
    per launch over the mean launch time, both summed over every profiled solve."""
    kernel_ms = sum(r.kernel_ms for r in results)
    relabel_ms = sum(r.relabel_ms for r in results)
    timed = sum(r.profiled_launches for r in results)
    launches = sum(r.cycles for r in results)
    if kernel_ms >= relabel_ms and timed:
        kname = "k_cycle"
        per_launch_bytes = sum(r.bytes_algorithmic for r in results) / launches
        per_launch_ms = sum(r.profiled_kernel_ms for r in results) / timed
    else:
        kname = "k_bfs_wg"  # global relabel; its time is taken per BFS
        launches = sum(r.rounds for r in results)
        per_launch_bytes = sum(r.bytes_bfs for r in results) / max(launches, 1)
        per_launch_ms = relabel_ms / max(launches, 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": kernel_traffic(kname, workload),
            "bytes_per_launch": int(per_launch_bytes),
            "mean_launch_us": round(per_launch_ms * 1e3, 3), "launches": int(launches),
            "timed_launches": int(timed) if kname == "k_cycle" else int(launches)}


def line_common(args, value, unit, ms_per_step, config, roofline, cpu, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}


# ---------------------------------------------------------------------------------------------
def bench_single(args, mf):
    W = WORKLOADS[args.workload]
    g = getattr(mf, W["gen"][0])(*W["gen"][1])
    algo_name = args.algo or W["algo"]
    algo, layout = mf.ALGOS[algo_name], mf.LAYOUTS[args.layout]
    opts = dict(cycles_per_relabel=args.cycles) if args.cycles else {}
    ctx = mf.Context(0)
    dev_name, ncu = ctx.device_info()
    log(f"device {dev_name}, {ncu} CUs; workload {W['desc']}")
    t_up = time.perf_counter()
    dg = ctx.upload(g, layout)  # host layout build + host-to-device copy
    upload_ms = 1e3 * (time.perf_counter() - t_up)

    def solve_checked(**kw):
        r = dg.solve(algo, **{**opts, **kw})
        if r.value != W["value"]:
            raise SystemExit(f"max-flow value {r.value} != exact {W['value']}")
        return r

    for _ in range(args.warmup):
        solve_checked()
    ctx.synchronize()
    # launch parameters from the autotuner (trial solves of this graph), before the timed solves
    tuned, trials = {}, 0
    if not args.no_autotune:
        tuned, tune_ms, trials = dg.autotune(algo, **opts)
        tuned = {k: v for k, v in tuned.items() if k not in ("batch_algo", "batch_mode")}
        log(f"autotune: {trials} trial solves, best {tune_ms:.2f} ms with {tuned}")
    # every 16th cycle launch of the timed solves carries start/end events in its dispatch
    # packet (a timed launch costs a cache flush, so only a sample is timed)
    t0 = time.perf_counter()
    results = [solve_checked(profile_kernels=PROFILE_EVERY, **tuned) for _ in range(args.steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    slots = 2 * g.m
    roofline = roofline_of(results, args.workload)
    t1 = time.perf_counter()
    plain = [solve_checked(**tuned) for _ in range(min(3, args.steps))]
    plain_ms = 1e3 * (time.perf_counter() - t1) / len(plain)
    t1 = time.perf_counter()
    dflt = [solve_checked() for _ in range(min(3, args.steps))]
    default_ms = 1e3 * (time.perf_counter() - t1) / len(dflt)
    log(f"{algo_name}/{args.layout}: {ms_per_step:.2f} ms per solve with launch events, "
        f"{plain_ms:.2f} ms without, {default_ms:.2f} ms with default options")

    variants = {}
    if not args.no_variants:
        # thread-centric scanning leaves one lane per hub vertex on power-law graphs (minutes
        # per solve at R-MAT scale 22), so only the vertex-centric variants run there
        algos = W.get("variants", ("vc",) if W["algo"] == "vc" else ("tc", "vc"))
        for va in algos:
            for vl in ("rcsr", "bcsr", "bcsr2"):
                vg = ctx.upload(g, mf.LAYOUTS[vl])
                times = []
                for _ in range(2):
                    r = vg.solve(mf.ALGOS[va], **opts)
                    if r.value != W["value"]:
                        raise SystemExit(f"{va}/{vl}: value {r.value} != exact {W['value']}")
                    times.append(r.solve_ms)
                vg.close()
                variants[f"{va}/{vl}"] = round(min(times), 3)
                log(f"variant {va}/{vl}: best of 2 solves {variants[f'{va}/{vl}']} ms")
    cpu = None if args.no_cpu_baseline else cpu_baseline_reference(mf)
    last = results[-1]
    config = {"workload": W["desc"], "algo": algo_name, "layout": args.layout,
              "max_flow": W["value"], "residual_slots": slots, "global_relabels": last.rounds,
              "cycles": last.cycles, "pushes": last.pushes, "relabels": last.relabels,
              "slots_scanned": last.slots_scanned, "upload_ms": round(upload_ms, 2),
              "autotuned_options": tuned, "autotune_trials": trials,
              "untimed_ms_per_solve": round(plain_ms, 3),
              "default_options_ms_per_solve": round(default_ms, 3),
              "host_inclusive_residual_edges_per_s": round(slots / ((upload_ms + ms_per_step) * 1e-3), 1),
              "scanned_slots_per_s": round(last.slots_scanned / (last.solve_ms * 1e-3), 1),
              "device": dev_name, "variants_solve_ms": variants}
    return line_common(args, round(slots / (wall / args.steps), 1), "residual edges/s",
                       ms_per_step, config, roofline, cpu)


# ---------------------------------------------------------------------------------------------
def batched_mix(mf, k, seed, lo=256, hi=65536):
    rng = np.random.default_rng(seed)
    graphs = []
    for i in range(k):
        L = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        R = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        graphs.append(mf.gen_bipartite(L, R, 8.0, seed=splitmix(seed * 1000003 + i)))
    return graphs


def cpu_baseline_batched(graphs, values, sample):
    """CPU oracle (Dinic), one instance per thread over all host cores, on a sample."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc

    idx = list(range(0, len(graphs), max(1, len(graphs) // sample)))[:sample]
    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        got = list(ex.map(lambda i: orc.dinic_maxflow(graphs[i]), idx))
    dt = time.perf_counter() - t0
    bad = [i for i, v in zip(idx, got) if v != values[i]]
    if bad:
        raise SystemExit(f"batched values differ from the CPU oracle on instances {bad[:5]}")
    return {"value": round(len(idx) / dt, 3), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (Dinic, C) on {len(idx)} instances of the same mix "
                      f"(every {len(graphs) // len(idx)}th), one instance per thread, "
                      f"{cores} threads, {dt:.2f} s"}


def bench_batched(args, mf):
    k = args.instances
    t0 = time.perf_counter()
    graphs = batched_mix(mf, k, args.seed)
    nv = sum(g.n for g in graphs)
    na = sum(g.m for g in graphs)
    log(f"generated {k} bipartite instances: {nv} vertices, {na} arcs "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = mf.Context(0)
    dev_name, ncu = ctx.device_info()
    algo_mode = 1 if args.algo == "vc" else 0  # batch_algo: 0 TC, 1 VC
    tuned, trials = {"batch_algo": algo_mode}, 0
    if not args.no_autotune:
        t1 = time.perf_counter()
        tuned, tune_ms, trials = ctx.batch_autotune(graphs, sample=args.tune_sample,
                                                    batch_algo=algo_mode)
        log(f"batch autotune: {trials} trial batches of {args.tune_sample} instances in "
            f"{time.perf_counter() - t1:.1f} s, best {tune_ms:.2f} ms with {tuned}")
        algo_mode = tuned["batch_algo"]
    runs = []
    values = None
    for step in range(args.warmup + args.steps):
        t1 = time.perf_counter()
        vals, res = ctx.batch_solve(graphs, **tuned)
        wall = time.perf_counter() - t1
        if values is not None and not np.array_equal(values, vals):
            raise SystemExit("batched values changed between runs")
        values = vals
        log(f"batched solve: device {res.solve_ms:.1f} ms, call {1e3 * wall:.1f} ms, "
            f"rounds {res.rounds}, cycles {res.cycles}")
        if step >= args.warmup:
            runs.append(res)
    solve_ms = float(np.mean([r.solve_ms for r in runs]))
    batched_sps = k / (solve_ms * 1e-3)
    # one by one on a sample of the same mix, graphs resident, same kernels
    rng = np.random.default_rng(args.seed + 1)
    sample = sorted(rng.choice(k, size=min(args.one_by_one, k), replace=False).tolist())
    per = []
    for i in sample:
        dg = ctx.upload(graphs[i], mf.LAYOUT_BCSR)
        r = dg.solve(mf.ALGO_VC if algo_mode else mf.ALGO_TC)
        if r.value != values[i]:
            raise SystemExit(f"instance {i}: alone {r.value} != batched {values[i]}")
        per.append(r.solve_ms)
        dg.close()
    one_sps = 1e3 / float(np.mean(per))
    log(f"batched {batched_sps:.1f} solves/s, one-by-one {one_sps:.1f} solves/s "
        f"(mean {np.mean(per):.2f} ms over {len(per)} instances): x{batched_sps / one_sps:.2f}")
    cpu = None if args.no_cpu_baseline else cpu_baseline_batched(graphs, values, args.cpu_sample)
    config = {"workload": f"{k} bipartite matching networks, |L|,|R| log-uniform [256,65536], "
                          f"Poisson(8) arcs per left vertex, unit capacities, seed {args.seed}",
              "vertices": nv, "arcs": na, "algo": "vc" if algo_mode else "tc", "layout": "bcsr",
              "batched_solves_per_s": round(batched_sps, 2),
              "one_by_one_solves_per_s": round(one_sps, 2),
              "one_by_one_sample": len(per), "speedup_vs_one_by_one": round(batched_sps / one_sps, 2),
              "autotuned_options": tuned, "autotune_trials": trials,
              "autotune_sample": args.tune_sample, "device": dev_name}
    return line_common(args, round(batched_sps, 2), "solves/s", solve_ms, config,
                       None, cpu)


# ---------------------------------------------------------------------------------------------
def bench_copy(args):
    """hipMemcpy device-to-device of 1 GiB: one read and one write of every byte."""
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                   "libamdhip64.so"))
    nbytes = 1 << 30

    def ok(rc, what):
        if rc != 0:
            raise SystemExit(f"{what} failed with HIP error {rc}")

    ok(hip.hipSetDevice(0), "hipSetDevice")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    ok(hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)), "hipMemset")
    ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    t0 = 0.0
    for step in range(args.warmup + args.steps):
        if step == args.warmup:
            ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
            t0 = time.perf_counter()
        ok(hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3), "hipMemcpy")  # device to device
    ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    wall = time.perf_counter() - t0
    hip.hipFree(src)
    hip.hipFree(dst)
    gbs = 2 * nbytes * args.steps / wall / 1e9
    log(f"copy: {gbs:.1f} GB/s over {args.steps} copies of {nbytes} bytes")
    config = {"workload": f"device-to-device copy of {nbytes} bytes (read + write)",
              "bytes_per_copy": nbytes}
    return line_common(args, round(gbs, 1), "GB/s", 1e3 * wall / args.steps, config,
                       {"bound": "hbm", "kernel": "runtime copy", "achieved": round(gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "traffic": 2 * nbytes}, None)


def serving_requests(mf, n, seed):
    """Mixed requests, 1e3..1e6 vertices log-uniform, generator chosen uniformly."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        nv = math.exp(rng.uniform(math.log(1e3), math.log(1e6)))
        kind = int(rng.integers(0, 3))
        s = splitmix(seed * 7919 + i)
        if kind == 0:
            side = max(3, int(round(math.sqrt(nv))))
            g = mf.gen_washington_rlg(side, side, 10000, s)
        elif kind == 1:
            a = max(2, int(round(nv ** (1 / 3))))
            g = mf.gen_genrmf(a, a, 1, 10000, s)
        else:
            half = max(2, int(round(nv / 2)))
            g = mf.gen_bipartite(half, half, 8.0, s)
        reqs.append(g)
    return reqs


def bench_serving(args, mf):
    reqs = serving_requests(mf, args.requests, args.seed)
    rng = np.random.default_rng(args.seed + 5)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(reqs)))
    ctx = mf.Context(0)
    dev_name, _ = ctx.device_info()
    srv = mf.Server(ctx, hbm_budget_bytes=args.budget_gb << 30, max_inflight_launches=args.inflight)
    done_at = {}
    ids = {}
    t0 = time.perf_counter()
    nxt = 0
    while len(done_at) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            ids[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        srv.progress()
        for rid, i in list(ids.items()):
            r = srv.result(rid)
            if r is not None:
                done_at[i] = (time.perf_counter() - t0, r[0])
                del ids[rid]
        if nxt < len(reqs) and not ids:
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - (time.perf_counter() - t0))))
        else:
            time.sleep(0.0005)
    total = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    lat = np.array([done_at[i][0] - arrivals[i] for i in range(len(reqs))]) * 1e3
    # each served value equals the same request solved alone
    check = list(range(0, len(reqs), max(1, len(reqs) // 16)))
    for i in check:
        alone = ctx.solve(reqs[i], mf.ALGO_VC).value
        if alone != done_at[i][1]:
            raise SystemExit(f"request {i}: served {done_at[i][1]} != alone {alone}")
    sps = len(reqs) / total
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    log(f"served {len(reqs)} requests in {total:.2f} s: {sps:.2f} solves/s, "
        f"p50 {p50:.1f} ms, p99 {p99:.1f} ms, launches {stats['launches']}")
    config = {"workload": f"{len(reqs)} requests, Poisson arrivals at {args.rate}/s, mix of RLG / "
                          f"Genrmf / bipartite with 1e3..1e6 vertices (log-uniform), seed {args.seed}",
              "p50_latency_ms": round(p50, 2), "p99_latency_ms": round(p99, 2),
              "solves_per_s": round(sps, 3), "offered_load_per_s": args.rate,
              "inflight_launches": args.inflight, "launches": stats["launches"],
              "arena_peak_bytes": stats["arena_peak_bytes"], "checked_alone": len(check),
              "device": dev_name}
    return line_common(args, round(sps, 3), "solves/s", 1e3 * total / len(reqs), config,
                       None, None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genrmf64",
                    choices=sorted(WORKLOADS) + ["batched", "serving", "copy"])
    ap.add_argument("--algo", default=None, choices=["tc", "vc"])
    ap.add_argument("--layout", default="bcsr", choices=["bcsr", "rcsr", "bcsr2"])
    ap.add_argument("--cycles", type=int, default=0, help="push-relabel cycles per global relabel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default launch parameters")
    ap.add_argument("--tune-sample", type=int, default=1024,
                    help="batched: autotune sample (enough instances to fill every workgroup)")
    ap.add_argument("--instances", type=int, default=4096, help="batched: number of instances")
    ap.add_argument("--one-by-one", type=int, default=256, help="batched: one-by-one sample size")
    ap.add_argument("--cpu-sample", type=int, default=64, help="batched: CPU baseline sample")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: arrivals per second")
    ap.add_argument("--inflight", type=int, default=4, help="serving: concurrent launches")
    ap.add_argument("--budget-gb", type=int, default=64, help="serving: HBM budget")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    import wbpr_amd as mf

    if args.workload == "batched":
        if args.steps == 10 and args.warmup == 2:  # defaults sized for the single solve
            args.steps, args.warmup = 1, 0
        line = bench_batched(args, mf)
    elif args.workload == "serving":
        line = bench_serving(args, mf)
    else:
        line = bench_single(args, mf)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
