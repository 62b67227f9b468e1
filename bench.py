#!/usr/bin/env python3
"""Max-flow benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload genrmf64] [--algo tc|vc]
                    [--layout bcsr|rcsr|bcsr2] [--no-cpu-baseline] [--no-variants]

A step is one complete push-relabel max-flow solve of the workload graph, which stays resident
in HBM (every solve restarts from the capacities on the device).  Every solve's value must equal
the committed exact value of the workload; the run fails otherwise.

Prints ONE JSON line: value = residual edges per second = residual slots (2m) / solve wall time,
ms_per_step = max-flow wall time, plus
  roofline      algorithmic bytes per launch of the dominant kernel over its mean launch time
                (HIP events around every launch of that kernel, on the solver's stream)
  cpu_baseline  the reference's own serial solver (maxflow-serial, built from its sources into
                oracle/_ref/) on a bounded sample of the same generator family, pinned to one core
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "max-flow wall time + residual edges/s (exact value); batched solves/s vs one-by-one"
HBM_PEAK_GBS = 8000.0

# Workloads with their exact max-flow values (tests/golden/values.json holds the same numbers,
# each computed by the CPU oracles).
WORKLOADS = {
    "genrmf64": dict(
        desc="Genrmf a=64 b=64 c1=1 c2=10000 seed=1 (262,144 vertices, 1,290,240 arcs)",
        gen=("gen_genrmf", (64, 64, 1, 10000, 1)), value=19996127),
    "rlg64": dict(
        desc="Washington RLG 64x64 r=10000 seed=1 (4,098 vertices, 12,224 arcs)",
        gen=("gen_washington_rlg", (64, 64, 10000, 1)), value=466088),
}
CPU_SAMPLE = dict(desc="Genrmf a=14 b=14 c1=1 c2=10000 seed=1 (2,744 vertices, 12,740 arcs)",
                  gen=("gen_genrmf", (14, 14, 1, 10000, 1)), value=919190)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(mf):
    """Reference serial solver on the bounded sample, one core; residual edges per second."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genrmf64", choices=sorted(WORKLOADS))
    ap.add_argument("--algo", default="tc", choices=["tc", "vc"])
    ap.add_argument("--layout", default="bcsr", choices=["bcsr", "rcsr", "bcsr2"])
    ap.add_argument("--cycles", type=int, default=0, help="push-relabel cycles per global relabel")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import wbpr_amd as mf

    W = WORKLOADS[args.workload]
    g = getattr(mf, W["gen"][0])(*W["gen"][1])
    algo = mf.ALGOS[args.algo]
    layout = mf.LAYOUTS[args.layout]
    opts = dict(cycles_per_relabel=args.cycles) if args.cycles else {}
    ctx = mf.Context(0)
    dev_name, ncu = ctx.device_info()
    log(f"device {dev_name}, {ncu} CUs; workload {W['desc']}")
    dg = ctx.upload(g, layout)

    def solve_checked(**kw):
        r = dg.solve(algo, **{**opts, **kw})
        if r.value != W["value"]:
            raise SystemExit(f"max-flow value {r.value} != exact {W['value']}")
        return r

    for _ in range(args.warmup):
        solve_checked()
    ctx.synchronize()
    t0 = time.perf_counter()
    results = [solve_checked() for _ in range(args.steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    slots = 2 * g.m
    value = slots / (wall / args.steps)
    log(f"{args.algo}/{args.layout}: {ms_per_step:.2f} ms per solve "
        f"(device solve {sum(r.solve_ms for r in results) / len(results):.2f} ms)")

    # roofline of the dominant kernel, from a solve with HIP events around every launch
    rp = solve_checked(profile_kernels=1)
    cyc_ms, bfs_ms = rp.kernel_ms, rp.relabel_ms
    if cyc_ms >= bfs_ms and rp.profiled_launches:
        kname = "k_cycle (push-relabel cycle)"
        per_launch_bytes = rp.bytes_algorithmic / rp.profiled_launches
        per_launch_ms = rp.profiled_kernel_ms / rp.profiled_launches
        launches = rp.profiled_launches
    else:
        kname = "k_bfs_wg + k_bfs_level (global relabel)"
        per_launch_bytes = rp.bytes_bfs / max(rp.rounds, 1)
        per_launch_ms = bfs_ms / max(rp.rounds, 1)
        launches = rp.rounds
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "bytes_per_launch": int(per_launch_bytes),
                "mean_launch_us": round(per_launch_ms * 1e3, 3), "launches": int(launches)}

    variants = {}
    if not args.no_variants:
        for vname, (va, vl) in {"tc/rcsr": ("tc", "rcsr"), "vc/rcsr": ("vc", "rcsr"),
                                "tc/bcsr": ("tc", "bcsr"), "vc/bcsr": ("vc", "bcsr"),
                                "tc/bcsr2": ("tc", "bcsr2"), "vc/bcsr2": ("vc", "bcsr2")}.items():
            vg = ctx.upload(g, mf.LAYOUTS[vl])
            times = []
            for _ in range(3):
                r = vg.solve(mf.ALGOS[va], **opts)
                if r.value != W["value"]:
                    raise SystemExit(f"{vname}: value {r.value} != exact {W['value']}")
                times.append(r.solve_ms)
            vg.close()
            variants[vname] = round(min(times), 3)
        log(f"variants (best of 3 solve ms): {variants}")

    cpu = None if args.no_cpu_baseline else cpu_baseline(mf)
    last = results[-1]
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "residual edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": W["desc"], "algo": args.algo, "layout": args.layout,
                   "max_flow": W["value"], "residual_slots": slots,
                   "global_relabels": last.rounds, "cycles": last.cycles,
                   "pushes": last.pushes, "relabels": last.relabels,
                   "device": dev_name, "variants_solve_ms": variants},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
