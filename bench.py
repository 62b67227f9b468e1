#!/usr/bin/env python3
"""Max-flow benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--algo tc|vc]
                    [--layout bcsr|rcsr|bcsr2] [--autotune | --default-options | --options JSON]
                    [--no-graphs] [--no-batched] [--no-cpu-baseline] [--no-variants]

Workloads
  rmat22 (default)    R-MAT scale 22 (4,194,304 vertices, 65.2M arcs), VC on BCSR: a step is one
                      complete max-flow solve of the resident graph (each solve restarts from the
                      capacities on the device).  The same JSON line also carries the batched
                      figures (4096 bipartite instances solved in one batched call against the
                      same instances solved one by one) unless --no-batched
  genrmf64, genrmf128, rlg64
                      the other single-instance configurations
  batched             the 4096-instance batched configuration alone (value = batched solves/s)
  serving             Poisson arrivals of mixed requests (RLG, Genrmf, bipartite; 1e3..1e6
                      vertices) through the continuous-batching server; p50/p99 latency
  copy                1 GiB device-to-device copies: the achievable HBM rate, and the known-size
                      kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters

Launch options of the single-instance solves: by default the ones committed in TUNED below
(found by this script's --autotune on an MI355X), so every run -- and the counter profiles
under profiles/ -- use the same options; --autotune runs the autotuner on the resident graph
first, --default-options uses the library defaults, --options sets them explicitly.
--no-graphs turns captured cycle chunks off (the counter runs, whose profiler cannot replay
graphs; the kernels and their arguments are the same).

Every value is checked exactly: single instances against their committed values, batched and
served results against solves of the same instances one by one.  Prints ONE JSON line (the
driver's contract) with a `roofline` entry for the dominant kernel (algorithmic bytes per launch
over the mean launch time, from HIP events on the solver's stream) and a `cpu_baseline` entry
(the CPU oracle and the reference's own serial solver, on the same graph, one pinned core).
"""
import argparse
import ctypes
import json
import os
import shutil
import signal
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
    "rmat22": dict(
        desc="R-MAT scale 22 edge factor 16 a=.57 b=.19 c=.19 caps [1,1000] seed=1 "
             "(4,194,304 vertices, 65,244,556 arcs after dropping self-loops and duplicates)",
        gen=("gen_rmat", (22, 16, 0.57, 0.19, 0.19, 1000, 1)), value=18886759, algo="vc"),
    "genrmf64": dict(
        desc="Genrmf a=64 b=64 c1=1 c2=10000 seed=1 (262,144 vertices, 1,290,240 arcs)",
        gen=("gen_genrmf", (64, 64, 1, 10000, 1)), value=19996127, algo="vc",
        variants=("tc", "vc")),
    "genrmf128": dict(
        desc="Genrmf a=128 b=128 c1=1 c2=10000 seed=1 (2,097,152 vertices, 10,403,840 arcs)",
        gen=("gen_genrmf", (128, 128, 1, 10000, 1)), value=81048332, algo="vc",
        variants=("tc", "vc")),
    "rlg64": dict(
        desc="Washington RLG 64x64 r=10000 seed=1 (4,098 vertices, 12,224 arcs)",
        gen=("gen_washington_rlg", (64, 64, 10000, 1)), value=466088, algo="tc"),
}

# Launch options per workload and algorithm, found by
#   python bench.py --workload <w> --autotune --no-batched --no-variants --no-cpu-baseline
# on one MI355X (the autotuner's coordinate descent over trial solves of the resident graph).
# R-MAT 22: two autotuner runs picked these options with a fixed cadence of 12 cycles per global
# relabel (8.3 ms) and with the adaptive cadence (7.97 ms); over 9 solves each the two ran 8.0 and
# 7.95 ms.  The fixed cadence is committed: it launches the same cycles whether or not they are
# replayed from captured graphs, so the counter runs (which cannot replay graphs) measure the
# launches the timed run makes; cooperative BFS levels cost R-MAT 0.16 ms (7.50 vs 7.66 ms).
# Genrmf 64: 156 trials, best 98.2 ms; then cooperative BFS levels for mid-size frontiers with the
# one-workgroup kernel's limit lowered to 256 vertices: 93.9 against 99.2 ms (4 solves each,
# interleaved).  Started from these options, the autotuner kept them on both graphs (R-MAT 22: 138
# trials, best 7.42 ms; Genrmf 64: 147 trials, best 93.7 ms).  Genrmf 64's tile widths then
# narrowed to 4 / 4 / 8 / 16 lanes up to 1,024 slots: 92.0 against 94.7 ms (8 interleaved solves
# each).  Workloads not listed run the library defaults.
TUNED = {
    ("rmat22", "vc"): {"cycles_per_relabel": 12, "tile_lanes": 0, "bfs_block_frontier": 1024,
                       "relabel_time_ratio": 2.0, "bfs_direction": 0,
                       "class_lanes": [4, 8, 16, 64, 256], "forward_hops": -1, "cycle_graphs": 1,
                       "bfs_cooperative": 0},
    ("genrmf64", "vc"): {"cycles_per_relabel": 674, "tile_lanes": 0, "bfs_block_frontier": 256,
                         "relabel_time_ratio": 4.0, "bfs_direction": 0,
                         "class_lanes": [4, 4, 8, 16, 256], "forward_hops": 2, "cycle_graphs": 0,
                         "bfs_cooperative": 1},
    # Genrmf 128: the library defaults plus cooperative BFS levels from 256 frontier vertices on
    # (559 against 708 ms without them; the autotuner, started there, kept them: 135 trials)
    ("genrmf128", "vc"): {"cycles_per_relabel": 0, "tile_lanes": 0, "bfs_block_frontier": 256,
                          "relabel_time_ratio": 2.0, "bfs_direction": 0,
                          "class_lanes": [8, 16, 32, 64, 256], "forward_hops": 2, "cycle_graphs": 0,
                          "bfs_cooperative": 1},
}

BATCH_K = 4096
BATCH_SEED = 2024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """Host CPU model and the cores this process may run on (the GPU box's share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cores = list(range(os.cpu_count() or 1))
    return {"cpu_model": model, "nproc": len(cores), "os_cpu_count": os.cpu_count()}, cores


class pinned:
    """Pins the calling thread to one core for the duration of a CPU-baseline leg."""

    def __init__(self, core):
        self.core = core

    def __enter__(self):
        self.old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {self.core})

    def __exit__(self, *a):
        os.sched_setaffinity(0, self.old)


def cpu_dinic_same_graph(g, expect, core):
    """The CPU oracle (Dinic, C, one thread pinned to one core) on the benchmarked graph."""
    from oracle import oracle as orc

    with pinned(core):
        t0 = time.perf_counter()
        v = orc.dinic_maxflow(g)
        dt = time.perf_counter() - t0
    if v != expect:
        raise SystemExit(f"cpu baseline: Dinic value {v} != exact {expect}")
    return dt


def cpu_reference_serial(g, core, cap_s):
    """The reference maxflow-serial (built from its sources into oracle/_ref) on the same graph,
    written as a DIMACS file, under a wall-clock cap.  Returns a dict; `seconds` is None when it
    did not finish (reported as "> cap")."""
    from oracle import oracle as orc

    if not orc.ref_available():
        return {"available": False}
    base = "/dev/shm" if os.path.isdir("/dev/shm") else None
    tmp = tempfile.mkdtemp(dir=base)
    path = os.path.join(tmp, "graph.max")
    try:
        t0 = time.perf_counter()
        g.save_dimacs(path)
        write_s = time.perf_counter() - t0
        cmd = [orc.REF_BIN, "-v", "2", "-f", path, "-s", "0", "-t", "1"]
        if shutil.which("taskset"):
            cmd = ["taskset", "-c", str(core)] + cmd
        t0 = time.perf_counter()
        proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                                start_new_session=True)
        try:
            out, _ = proc.communicate(timeout=cap_s)
            dt = time.perf_counter() - t0
            got = None
            for line in out.splitlines():
                if line.startswith("Max flow:"):
                    got = int(line.split()[-1])
            return {"available": True, "cap_s": cap_s, "seconds": round(dt, 2), "value": got,
                    "dimacs_write_s": round(write_s, 2)}
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.communicate()
            return {"available": True, "cap_s": cap_s, "seconds": None,
                    "dimacs_write_s": round(write_s, 2)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


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
    """Roofline of the dominant kernel over solves run with profile_kernels: algorithmic bytes
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


def device_label(ctx):
    name, ncu = ctx.device_info()
    label = name.strip()
    if label.startswith("("):  # the runtime reported no marketing name
        label = label.strip("()").split(":")[0]
    return f"{label}, {ncu} CUs", ncu


def solve_options(args, workload, algo_name):
    """Options of the timed solves (see the module docstring) and where they came from."""
    if args.options:
        opts, src = json.loads(args.options), "--options"
    elif args.default_options:
        opts, src = {}, "library defaults"
    else:
        opts = dict(TUNED.get((workload, algo_name), {}))
        src = "committed tuned options" if opts else "library defaults (none committed)"
    if args.cycles:
        opts["cycles_per_relabel"] = args.cycles
    if args.no_graphs:
        opts["cycle_graphs"] = 0
    return opts, src


# ---------------------------------------------------------------------------------------------
def bench_single(args, mf):
    W = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    g = getattr(mf, W["gen"][0])(*W["gen"][1])
    gen_s = time.perf_counter() - t0
    algo_name = args.algo or W["algo"]
    algo, layout = mf.ALGOS[algo_name], mf.LAYOUTS[args.layout]
    ctx = mf.Context(0)
    dev, _ = device_label(ctx)
    log(f"device {dev}; workload {W['desc']} (generated in {gen_s:.1f} s)")
    t_up = time.perf_counter()
    dg = ctx.upload(g, layout)  # device layout build from the host CSR + host-to-device copy
    upload_ms = 1e3 * (time.perf_counter() - t_up)

    def solve_checked(**kw):
        r = dg.solve(algo, **kw)
        if r.value != W["value"]:
            raise SystemExit(f"max-flow value {r.value} != exact {W['value']}")
        return r

    opts, opts_src = solve_options(args, args.workload, algo_name)
    trials = 0
    if args.autotune:
        tuned, tune_ms, trials = dg.autotune(algo, **opts)
        opts = {k: v for k, v in tuned.items() if k not in ("batch_algo", "batch_mode")}
        if args.no_graphs:
            opts["cycle_graphs"] = 0
        opts_src = f"autotuned in this run ({trials} trial solves)"
        log(f"autotune: {trials} trial solves, best {tune_ms:.2f} ms with {json.dumps(opts)}")
    for _ in range(args.warmup):
        solve_checked(**opts)
    ctx.synchronize()
    # every 16th cycle launch of the timed solves carries start/end events in its dispatch
    # packet (a timed launch costs a cache flush, so only a sample is timed)
    t0 = time.perf_counter()
    results = [solve_checked(profile_kernels=PROFILE_EVERY, **opts) for _ in range(args.steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    slots = 2 * g.m
    roofline = roofline_of(results, args.workload)
    solve_ms = [r.solve_ms for r in results]
    t1 = time.perf_counter()
    plain = [solve_checked(**opts) for _ in range(min(3, args.steps))]
    plain_ms = 1e3 * (time.perf_counter() - t1) / len(plain)
    log(f"{algo_name}/{args.layout}: {ms_per_step:.2f} ms per solve with launch events "
        f"(min {min(solve_ms):.2f}, max {max(solve_ms):.2f}), {plain_ms:.2f} ms without")
    variants = {}
    if not args.no_variants:
        # thread-centric scanning leaves one lane per hub vertex on power-law graphs (minutes
        # per solve at R-MAT scale 22), so only the vertex-centric variants run there
        algos = W.get("variants", ("vc",) if W["algo"] == "vc" else ("tc", "vc"))
        for va in algos:
            for vl in ("rcsr", "bcsr", "bcsr2"):
                vg = ctx.upload(g, mf.LAYOUTS[vl])
                times = []
                for _ in range(3):
                    r = vg.solve(mf.ALGOS[va])
                    if r.value != W["value"]:
                        raise SystemExit(f"{va}/{vl}: value {r.value} != exact {W['value']}")
                    times.append(r.solve_ms)
                vg.close()
                variants[f"{va}/{vl}"] = round(min(times), 3)
                log(f"variant {va}/{vl} (default options): best of 3 solves "
                    f"{variants[f'{va}/{vl}']} ms")
    dg.close()
    last = results[-1]
    config = {"workload": W["desc"], "algo": algo_name, "layout": args.layout,
              "max_flow": W["value"], "value_is": "residual slots of the graph (2m) per second "
              "of solve wall time, graph resident in HBM",
              "residual_slots": slots, "solve_ms_mean": round(ms_per_step, 3),
              "solve_ms_min": round(min(solve_ms), 3), "solve_ms_max": round(max(solve_ms), 3),
              "global_relabels": last.rounds, "cycles": last.cycles, "pushes": last.pushes,
              "relabels": last.relabels, "slots_scanned": last.slots_scanned,
              "scanned_slots_per_s": round(last.slots_scanned / (last.solve_ms * 1e-3), 1),
              "kernel_ms": round(last.kernel_ms, 3), "relabel_ms": round(last.relabel_ms, 3),
              "upload_ms": round(upload_ms, 2),
              "host_inclusive_residual_edges_per_s":
                  round(slots / ((upload_ms + ms_per_step) * 1e-3), 1),
              "options": opts, "options_source": opts_src, "autotune_trials": trials,
              "untimed_ms_per_solve": round(plain_ms, 3), "device": dev,
              "variants_solve_ms": variants}
    value = round(slots / (wall / args.steps), 1)
    cpu = None
    if not args.no_cpu_baseline:
        info, cores = host_info()
        core = cores[0]
        log(f"cpu baseline: Dinic on the same graph, pinned to core {core} ({info['cpu_model']})")
        dinic_s = cpu_dinic_same_graph(g, W["value"], core)
        log(f"cpu baseline: Dinic {dinic_s:.2f} s; reference serial under a {args.ref_cap} s cap")
        ref = cpu_reference_serial(g, core, args.ref_cap)
        if not ref.get("available"):
            ref_txt = "the reference maxflow-serial binary is not built"
        elif ref["seconds"] is None:
            ref_txt = (f"the reference maxflow-serial (built from its sources) on the same graph "
                       f"as DIMACS did not finish within its {args.ref_cap} s cap (> {args.ref_cap} s)")
        else:
            ref_txt = (f"the reference maxflow-serial on the same graph: {ref['seconds']} s "
                       f"(value {ref['value']})")
        cpu = {"value": round(slots / dinic_s, 1), "unit": "residual edges/s", "cores": 1,
               "kind": "port",
               "sample": f"the whole benchmarked graph ({W['desc']}): CPU oracle (Dinic, C, one "
                         f"thread pinned to core {core}) {dinic_s:.2f} s per solve, value = 2m / "
                         f"solve time; {ref_txt}",
               "seconds": round(dinic_s, 3), "host": info, "reference_serial": ref}
        log(f"cpu baseline: {ref_txt}")
    return line_common(args, value, "residual edges/s", ms_per_step, config, roofline, cpu)


# ---------------------------------------------------------------------------------------------
def cpu_baseline_batched(graphs, values, sample, cores):
    """CPU oracle (Dinic), one instance per thread over the host cores, on a sample."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc

    idx = list(range(0, len(graphs), max(1, len(graphs) // sample)))[:sample]
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


def run_batched(args, mf):
    """The 4096-instance batched configuration: one batched call against the same instances
    solved one by one with the single-instance kernels.  Returns (summary dict, mean ms)."""
    from wbpr_amd.workloads import batched_mix

    k = args.instances
    t0 = time.perf_counter()
    graphs = batched_mix(mf, k, args.seed)
    nv = sum(g.n for g in graphs)
    na = sum(g.m for g in graphs)
    log(f"generated {k} bipartite instances: {nv} vertices, {na} arcs "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = mf.Context(0)
    dev, _ = device_label(ctx)
    algo_mode = 1 if args.algo == "vc" else 0  # batch_algo: 0 TC, 1 VC
    tuned, trials = {"batch_algo": algo_mode}, 0
    if not args.no_batch_autotune:
        t1 = time.perf_counter()
        tuned, tune_ms, trials = ctx.batch_autotune(graphs, sample=args.tune_sample,
                                                    batch_algo=algo_mode)
        log(f"batch autotune: {trials} trial batches of {args.tune_sample} instances in "
            f"{time.perf_counter() - t1:.1f} s, best {tune_ms:.2f} ms with {tuned}")
        algo_mode = tuned["batch_algo"]
    runs, calls = [], []
    values = None
    for step in range(args.batch_warmup + args.batch_steps):
        t1 = time.perf_counter()
        vals, res = ctx.batch_solve(graphs, **tuned)
        wall = time.perf_counter() - t1
        if values is not None and not np.array_equal(values, vals):
            raise SystemExit("batched values changed between runs")
        values = vals
        log(f"batched solve: device {res.solve_ms:.1f} ms, call {1e3 * wall:.1f} ms, "
            f"rounds {res.rounds}, cycles {res.cycles}")
        if step >= args.batch_warmup:
            runs.append(res)
            calls.append(1e3 * wall)
    solve_ms = float(np.mean([r.solve_ms for r in runs]))
    batched_sps = k / (solve_ms * 1e-3)
    # one by one on a random sample of the same mix, each graph resident, single-instance solver
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
    cpu = None
    if not args.no_cpu_baseline:
        info, cores = host_info()
        cpu = cpu_baseline_batched(graphs, values, args.cpu_sample, min(16, len(cores)))
        cpu["host"] = info
    summary = {"workload": f"{k} bipartite matching networks, |L|,|R| log-uniform [256,65536], "
                           f"Poisson(8) arcs per left vertex, unit capacities, seed {args.seed}",
               "vertices": nv, "arcs": na, "algo": "vc" if algo_mode else "tc",
               "layout": "bcsr", "batched_solves_per_s": round(batched_sps, 2),
               "batched_device_ms": round(solve_ms, 3),
               "batched_call_ms": round(float(np.mean(calls)), 1),
               "one_by_one_solves_per_s": round(one_sps, 2),
               "one_by_one_sample": len(per),
               "speedup_vs_one_by_one": round(batched_sps / one_sps, 2),
               "options": tuned, "autotune_trials": trials,
               "autotune_sample": args.tune_sample if trials else 0, "device": dev,
               "cpu_baseline": cpu}
    return summary, solve_ms


def bench_batched(args, mf):
    summary, solve_ms = run_batched(args, mf)
    cpu = summary.pop("cpu_baseline")
    args.steps, args.warmup = args.batch_steps, args.batch_warmup
    return line_common(args, summary["batched_solves_per_s"], "solves/s", solve_ms, summary,
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


def serve_schedule(mf, ctx, reqs, arrivals, budget_bytes, inflight):
    """Submits reqs at their arrival times (seconds from the start) to one mf.Server and polls
    until every one is done.  Returns (values, seconds, library latency, stats)."""
    srv = mf.Server(ctx, hbm_budget_bytes=budget_bytes, max_inflight_launches=inflight)
    done = {}
    ids = {}
    t0 = time.perf_counter()
    nxt = 0
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            ids[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        srv.progress()
        for rid in srv.completed():
            done[ids.pop(rid)] = srv.result(rid)[0]
        if nxt < len(reqs) and not ids:
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - (time.perf_counter() - t0))))
        else:
            time.sleep(0.0002)
    total = time.perf_counter() - t0
    stats = srv.stats()
    lat = srv.latency()  # (p50 ms, p99 ms, solves/s) of finish - submit on the library's clock
    srv.close()
    return [done[i] for i in range(len(reqs))], total, lat, stats


def check_alone(mf, ctx, reqs, values, count=16):
    """Each served value of a spread sample equals the same request solved alone."""
    check = list(range(0, len(reqs), max(1, len(reqs) // count)))
    for i in check:
        alone = ctx.solve(reqs[i], mf.ALGO_VC).value
        if alone != values[i]:
            raise SystemExit(f"request {i}: served {values[i]} != alone {alone}")
    return len(check)


def bench_serving(args, mf):
    from wbpr_amd.workloads import poisson_arrivals, serving_requests

    ctx = mf.Context(0)
    dev, _ = device_label(ctx)
    budget = args.budget_gb << 30
    # the serving configuration: the three generators at 1e3..1e6 vertices, at the offered load
    reqs, kinds = serving_requests(mf, args.requests, args.seed)
    arrivals = poisson_arrivals(len(reqs), args.rate, args.seed + 5)
    values, total, (p50, p99, lib_sps), stats = serve_schedule(mf, ctx, reqs, arrivals, budget,
                                                                args.inflight)
    checked = check_alone(mf, ctx, reqs, values)
    sps = len(reqs) / total
    log(f"served {len(reqs)} requests in {total:.2f} s: {sps:.2f} solves/s, library latency "
        f"p50 {p50:.1f} ms, p99 {p99:.1f} ms; {stats['batched']} joined the serving kernel, "
        f"{stats['solo']} ran alone, {stats['serve_kernel_launches']} serving-kernel launches")
    # a high load of small requests (each joins the running serving kernel): the requests that
    # arrive while others are on the GPU share its launches
    breqs, _ = serving_requests(mf, args.burst_requests, args.seed + 11, vmin=1e3, vmax=1.6e4)
    barr = poisson_arrivals(len(breqs), args.burst_rate, args.seed + 13)
    bvalues, btotal, (bp50, bp99, bsps), bstats = serve_schedule(mf, ctx, breqs, barr, budget,
                                                                 args.inflight)
    bchecked = check_alone(mf, ctx, breqs, bvalues)
    log(f"burst: {len(breqs)} requests at {args.burst_rate}/s in {btotal:.2f} s: "
        f"{len(breqs) / btotal:.1f} solves/s, p50 {bp50:.1f} ms, p99 {bp99:.1f} ms; "
        f"{bstats['batched']} joined, {bstats['serve_kernel_launches']} serving-kernel launches")
    config = {"workload": f"{len(reqs)} requests, Poisson arrivals at {args.rate}/s, mix of RLG / "
                          f"Genrmf / bipartite with 1e3..1e6 vertices (log-uniform), seed {args.seed}",
              "p50_latency_ms": round(p50, 2), "p99_latency_ms": round(p99, 2),
              "latency_is": "mf_server_latency: finish - submit on the library's clock",
              "solves_per_s": round(sps, 3), "library_solves_per_s": round(lib_sps, 3),
              "offered_load_per_s": args.rate,
              "inflight_launches": args.inflight, "launches": stats["launches"],
              "joined_serving_kernel": stats["batched"], "ran_alone": stats["solo"],
              "serving_kernel_launches": stats["serve_kernel_launches"],
              "arena_peak_bytes": stats["arena_peak_bytes"], "checked_alone": checked,
              "high_load": {
                  "workload": f"{len(breqs)} requests, Poisson arrivals at {args.burst_rate}/s, "
                              f"same generators at 1e3..1.6e4 vertices, seed {args.seed + 11}",
                  "p50_latency_ms": round(bp50, 2), "p99_latency_ms": round(bp99, 2),
                  "solves_per_s": round(len(breqs) / btotal, 2),
                  "joined_serving_kernel": bstats["batched"], "ran_alone": bstats["solo"],
                  "serving_kernel_launches": bstats["serve_kernel_launches"],
                  "checked_alone": bchecked},
              "device": dev}
    return line_common(args, round(sps, 3), "solves/s", 1e3 * total / len(reqs), config,
                       None, None)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22",
                    choices=sorted(WORKLOADS) + ["batched", "serving", "copy"])
    ap.add_argument("--algo", default=None, choices=["tc", "vc"])
    ap.add_argument("--layout", default="bcsr", choices=["bcsr", "rcsr", "bcsr2"])
    ap.add_argument("--cycles", type=int, default=0, help="push-relabel cycles per global relabel")
    grp = ap.add_mutually_exclusive_group()
    grp.add_argument("--autotune", action="store_true",
                     help="run the autotuner on the resident graph before the timed solves")
    grp.add_argument("--default-options", action="store_true", help="library default options")
    grp.add_argument("--options", default=None, help="explicit options as a JSON object")
    ap.add_argument("--no-graphs", action="store_true",
                    help="no captured cycle chunks (counter runs)")
    ap.add_argument("--no-batched", action="store_true",
                    help="single-instance workloads: leave out the batched figures")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--ref-cap", type=float, default=30.0,
                    help="wall-clock cap (s) of the reference serial solver's run")
    ap.add_argument("--no-batch-autotune", action="store_true")
    ap.add_argument("--tune-sample", type=int, default=1024,
                    help="batched: autotune sample (enough instances to fill every workgroup)")
    ap.add_argument("--instances", type=int, default=BATCH_K, help="batched: number of instances")
    ap.add_argument("--batch-steps", type=int, default=2, help="batched: timed batched calls")
    ap.add_argument("--batch-warmup", type=int, default=1, help="batched: untimed batched calls")
    ap.add_argument("--one-by-one", type=int, default=512, help="batched: one-by-one sample size")
    ap.add_argument("--cpu-sample", type=int, default=64, help="batched: CPU baseline sample")
    ap.add_argument("--seed", type=int, default=BATCH_SEED)
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: arrivals per second")
    ap.add_argument("--inflight", type=int, default=4, help="serving: concurrent launches")
    ap.add_argument("--budget-gb", type=int, default=64, help="serving: HBM budget")
    ap.add_argument("--burst-requests", type=int, default=512,
                    help="serving: small requests of the high-load phase")
    ap.add_argument("--burst-rate", type=float, default=2000.0,
                    help="serving: arrivals per second of the high-load phase")
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
        line = bench_batched(args, mf)
    elif args.workload == "serving":
        line = bench_serving(args, mf)
    else:
        line = bench_single(args, mf)
        if not args.no_batched:
            summary, _ = run_batched(args, mf)
            line["config"]["batched"] = summary
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
