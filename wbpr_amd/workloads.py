"""Seeded workloads of the batched and serving configurations, shared by the tests and the
benchmark so both run exactly the same instances.

  batched_mix        k bipartite matching networks, |L| and |R| each log-uniform in [lo, hi],
                     Poisson(8) distinct arcs per left vertex, unit capacities (the paper's
                     bipartite formulation, WBPR_HPEC2025_CameraReady.pdf p.6 Table III)
  serving_requests   requests mixing the Washington RLG, Genrmf and bipartite generators with a
                     log-uniform vertex count in [vmin, vmax]
  poisson_arrivals   arrival times (seconds) of a Poisson process of a given rate

Each instance gets its own splitmix64-derived seed from the master seed, so instance i is the
same whatever k is.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = 0xFFFFFFFFFFFFFFFF


def splitmix(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def batched_sizes(k: int, seed: int, lo: int = 256, hi: int = 65536):
    """(|L|, |R|) of every instance of batched_mix, without generating the graphs."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        L = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        R = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        out.append((L, R))
    return out


def batched_mix(mf, k: int, seed: int, lo: int = 256, hi: int = 65536, mean_deg: float = 8.0):
    """k bipartite matching networks: source 0 -> L left -> R right -> sink, unit capacities."""
    return [mf.gen_bipartite(L, R, mean_deg, seed=splitmix(seed * 1000003 + i))
            for i, (L, R) in enumerate(batched_sizes(k, seed, lo, hi))]


def serving_requests(mf, n: int, seed: int, vmin: float = 1e3, vmax: float = 1e6):
    """n requests; each picks a generator uniformly and a vertex count log-uniform in
    [vmin, vmax]: RLG with sqrt(V) levels of sqrt(V) vertices, Genrmf with a = b = V^(1/3),
    bipartite with V/2 vertices per side.  Capacities: RLG and Genrmf [1, 10000], bipartite 1.
    Returns (graphs, kinds) with kinds[i] in {"rlg", "genrmf", "bipartite"}."""
    rng = np.random.default_rng(seed)
    graphs, kinds = [], []
    for i in range(n):
        nv = math.exp(rng.uniform(math.log(vmin), math.log(vmax)))
        kind = int(rng.integers(0, 3))
        s = splitmix(seed * 7919 + i)
        if kind == 0:
            side = max(3, int(round(math.sqrt(nv))))
            graphs.append(mf.gen_washington_rlg(side, side, 10000, s))
            kinds.append("rlg")
        elif kind == 1:
            a = max(2, int(round(nv ** (1 / 3))))
            graphs.append(mf.gen_genrmf(a, a, 1, 10000, s))
            kinds.append("genrmf")
        else:
            half = max(2, int(round(nv / 2)))
            graphs.append(mf.gen_bipartite(half, half, 8.0, s))
            kinds.append("bipartite")
    return graphs, kinds


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times in seconds of n requests of a Poisson process with `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
