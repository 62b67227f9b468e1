import json
import os
import sys

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
