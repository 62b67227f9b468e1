"""CPU oracle for the max-flow path.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, as the
checker; the product package (wbpr_amd) never imports it.

  serial_maxflow   C restatement of the reference CPU solver (maxflow-serial/graph.cpp:316-619)
  dinic_maxflow    independent exact max flow (Dinic), also returns a maximum flow
  certify          capacity / conservation / value / cut certificate of a flow
  brute_force_min_cut   minimum s-t cut by enumerating every vertex subset (n <= 18)
  ref_serial       runs the reference's own serial binary built by oracle/Makefile into _ref/
  ref_layouts      runs the reference's own loaders and residual-layout builders (its
                   maxflow-cuda / maxflow-bcsr / maxflow-bcsr2 src/graph.cpp, built with our
                   dump driver ref_layouts.cpp into _ref/) and returns their arrays
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "maxflow-serial")
REF_LAYOUTS = {k: os.path.join(HERE, "_ref", f"ref_layouts_{k}") for k in ("rcsr", "bcsr", "bcsr2")}

if not os.path.exists(LIB):
    raise ImportError(f"{LIB} missing: run `make -C oracle`")
_lib = C.CDLL(LIB)
_i32p = C.POINTER(C.c_int)
_lib.oracle_serial_maxflow.restype = C.c_longlong
_lib.oracle_serial_maxflow.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int, C.c_int,
                                       C.POINTER(C.c_longlong)]
_lib.oracle_dinic_maxflow.restype = C.c_longlong
_lib.oracle_dinic_maxflow.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int, C.c_int,
                                      _i32p]
_lib.oracle_certify.restype = C.c_int
_lib.oracle_certify.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, C.c_int, C.c_int, _i32p,
                                C.c_longlong, C.c_int, C.POINTER(C.c_longlong)]

_lib.oracle_rcsr_build.restype = None
_lib.oracle_rcsr_build.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
_lib.oracle_bcsr_build.restype = None
_lib.oracle_bcsr_build.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                   _i32p]

CERT_CODES = {0: "ok", 1: "capacity", 2: "conservation", 3: "value", 4: "cut"}


def rcsr_build(g):
    """Reference RCSR reverse arrays: (roffsets, rdst, flow_index)."""
    keep, po, pd, pc = _csr(g)
    roff = np.zeros(g.n + 1, dtype=np.int32)
    rdst = np.zeros(max(g.m, 1), dtype=np.int32)
    fidx = np.zeros(max(g.m, 1), dtype=np.int32)
    _lib.oracle_rcsr_build(g.n, g.m, po, pd, roff.ctypes.data_as(_i32p),
                           rdst.ctypes.data_as(_i32p), fidx.ctypes.data_as(_i32p))
    return roff, rdst[: g.m], fidx[: g.m]


def bcsr_build(n, arc_u, arc_v, arc_cap):
    """Reference BCSR from arcs in file order: (offsets, dst, residual, backward_idx)."""
    u, pu = _arr(arc_u)
    v, pv = _arr(arc_v)
    c, pc = _arr(arc_cap)
    m = len(u)
    off = np.zeros(n + 1, dtype=np.int32)
    out = [np.zeros(max(2 * m, 1), dtype=np.int32) for _ in range(3)]
    _lib.oracle_bcsr_build(n, m, pu, pv, pc, off.ctypes.data_as(_i32p),
                           *(o.ctypes.data_as(_i32p) for o in out))
    return (off,) + tuple(o[: 2 * m] for o in out)


def _arr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def _csr(g):
    off, po = _arr(g.offsets)
    dst, pd = _arr(g.dst)
    cap, pc = _arr(g.cap)
    return (off, dst, cap), po, pd, pc


def serial_maxflow(g, s=None, t=None):
    """Value as maxflow-serial computes it (None when it reports no s->t path)."""
    keep, po, pd, pc = _csr(g)
    ops = C.c_longlong()
    v = _lib.oracle_serial_maxflow(g.n, g.m, po, pd, pc, g.s if s is None else s,
                                   g.t if t is None else t, C.byref(ops))
    return None if v < 0 else int(v)


def dinic_maxflow(g, with_flow=False):
    keep, po, pd, pc = _csr(g)
    flow = np.zeros(max(g.m, 1), dtype=np.int32)
    v = _lib.oracle_dinic_maxflow(g.n, g.m, po, pd, pc, g.s, g.t, flow.ctypes.data_as(_i32p))
    return (int(v), flow[: g.m]) if with_flow else int(v)


def certify(g, flow, value, preflow=False):
    """Returns (code name, cut capacity)."""
    keep, po, pd, pc = _csr(g)
    fl, pf = _arr(flow)
    cut = C.c_longlong(-1)
    rc = _lib.oracle_certify(g.n, g.m, po, pd, pc, g.s, g.t, pf, int(value), int(preflow),
                             C.byref(cut))
    return CERT_CODES.get(rc, str(rc)), int(cut.value)


def brute_force_min_cut(g):
    """Minimum s-t cut over all vertex subsets (max-flow min-cut theorem); n <= 18."""
    n = g.n
    assert n <= 18
    tails = np.repeat(np.arange(n), np.diff(g.offsets))
    heads = np.asarray(g.dst)
    caps = np.asarray(g.cap, dtype=np.int64)
    others = [v for v in range(n) if v not in (g.s, g.t)]
    best = None
    for mask in range(1 << len(others)):
        side = np.zeros(n, dtype=bool)
        side[g.s] = True
        for b, v in enumerate(others):
            if mask >> b & 1:
                side[v] = True
        c = int(caps[side[tails] & ~side[heads]].sum())
        best = c if best is None else min(best, c)
    return best


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def ref_serial(path: str, file_type: int, s: int = 0, t: int = 1, timeout: float = 600):
    """Runs the reference serial binary (built from the reference sources) on a file.

    Returns (value or None when it prints no value, wall seconds)."""
    import time

    args = [REF_BIN, "-v", str(file_type), "-f", path, "-s", str(s), "-t", str(t)]
    t0 = time.perf_counter()
    out = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    dt = time.perf_counter() - t0
    m = re.search(r"Max flow: (-?\d+)", out.stdout)
    return (int(m.group(1)) if m else None), dt


def ref_layouts_available() -> bool:
    return all(os.path.exists(p) for p in REF_LAYOUTS.values())


def ref_layouts(layout: str, path: str, file_type: int, tmpdir: str) -> dict:
    """Arrays the reference's host code builds from a graph file (ref_layouts.cpp): its
    CSRGraph (csr_header = [n, m, m_processed, s, t], csr_offsets, csr_destinations,
    csr_capacities) and its ResidualGraph for `layout` ("rcsr": offsets, destinations,
    capacities, forward_flows, backward_flows, roffsets, rdestinations, flow_index; "bcsr": ...,
    forward_flows, backward_idx; "bcsr2": ..., forward_flows of 2m entries)."""
    out = os.path.join(tmpdir, f"ref_{layout}.txt")
    subprocess.run([REF_LAYOUTS[layout], str(file_type), path, out], check=True,
                   capture_output=True, timeout=600)
    arrays = {}
    with open(out) as f:
        for line in f:
            parts = line.split()
            n = int(parts[1])
            arrays[parts[0]] = np.array([int(x) for x in parts[2:2 + n]], dtype=np.int64)
    return arrays
