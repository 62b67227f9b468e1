import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def mf():
    import wbpr_amd

    return wbpr_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx(mf):
    c = mf.Context(0)
    yield c
    c.close()


def fixture_graph(mf, golden, name):
    rec = golden["fixtures"][name]
    g = mf.load(os.path.join(GOLDEN, name), rec["type"])
    if rec["type"] != 2:
        g = g.with_terminals(rec["s"], rec["t"])
    return g, rec["value"]


def generated_graph(mf, golden, name):
    rec = golden["generated"][name]
    return getattr(mf, rec["gen"])(*rec["args"]), rec["value"]


DEGENERATE = {  # (n, tails, heads, capacities, s, t)
    "no_arcs": (2, [], [], [], 0, 1),
    "parallel_s_t": (2, [0, 0, 0], [1, 1, 1], [1, 2, 3], 0, 1),
    "only_backwards_and_loops": (3, [1, 0, 1, 2], [0, 0, 1, 2], [4, 5, 6, 7], 0, 1),
    "loop_on_path": (4, [0, 1, 1, 1, 3], [1, 3, 1, 2, 1], [5, 3, 7, 1, 2], 0, 3),
    "zero_capacity": (3, [0, 0, 2], [1, 2, 1], [0, 0, 9], 0, 1),
    "isolated_terminals": (6, [2, 3, 4], [3, 4, 2], [1, 1, 1], 0, 5),
    # the largest capacities an upload accepts: the source's out-capacity at the int32 limit
    "int32_capacities": (4, [0, 1, 1, 2], [1, 3, 2, 3], [2**31 - 1, 2**31 - 6, 2**31 - 1, 3], 0, 3),
}


def degenerate_graph(mf, name):
    n, u, v, c, s, t = DEGENERATE[name]
    return mf.Graph.from_arcs(n, np.array(u, dtype=np.int64), np.array(v, dtype=np.int64),
                              np.array(c, dtype=np.int32), s, t)


def random_small_graph(mf, rng):
    n = int(rng.integers(2, 65))
    m = int(rng.integers(0, 6 * n + 1))
    u = rng.integers(0, n, size=m)
    v = rng.integers(0, n, size=m)
    s, t = (int(x) for x in rng.choice(n, size=2, replace=False))
    top = int(rng.choice([1, 10, 1000, 2**31 - 1]))
    cap = rng.integers(0, top, size=m, endpoint=True).astype(np.int64)
    out_s = np.nonzero(u == s)[0]
    if cap[out_s].sum() > 2**31 - 1:  # every excess must fit int32: scale the source arcs
        cap[out_s] = cap[out_s] // (len(out_s) + 1)
    return mf.Graph.from_arcs(n, u, v, cap.astype(np.int32), s, t)
