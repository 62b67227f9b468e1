"""Pins the CPU oracle: known answers of the reference's fixture files, agreement between the
restated serial solver, Dinic, the reference's own serial binary, scipy and brute-force cut
enumeration, and certificate checks that reject planted mistakes."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, fixture_graph, generated_graph

FIXTURE_VALUES = {  # known answers (published with the fixture files, reproduced by every solver)
    "test_DIMACS.txt": 25, "genrmf_a10_b1.max": 2000, "genrmf_a4_b3.max": 228, "ww.max": 77962,
    "m10.max": 9860177, "m30.max": 28258807, "m50.max": 42791871, "waissi_example.max": 6,
    "snap_8.txt": 4, "snap_5.txt": 3, "snap_4.txt": 2,
}


def random_graph(mf, rng, n, m, cap_max=20, parallel=False):
    u = rng.integers(0, n, size=m)
    v = rng.integers(0, n, size=m)
    keep = u != v
    u, v = u[keep], v[keep]
    if not parallel:
        pairs = np.unique(np.stack([u, v], 1), axis=0)
        u, v = pairs[:, 0], pairs[:, 1]
    cap = rng.integers(0, cap_max + 1, size=len(u))
    return mf.Graph.from_arcs(n, u, v, cap, 0, n - 1)


@pytest.mark.parametrize("name", sorted(FIXTURE_VALUES))
def test_fixture_known_answers(mf, orc, golden, name):
    g, value = fixture_graph(mf, golden, name)
    assert value == FIXTURE_VALUES[name]
    v, flow = orc.dinic_maxflow(g, with_flow=True)
    assert v == value
    assert orc.serial_maxflow(g) == value
    assert orc.certify(g, flow, value)[0] == "ok"


def test_genrmf_closed_form(mf, orc, golden):
    # Genrmf with b = 1 is one a x a frame: both corners have degree 2 and in-frame capacity
    # c2*a*a, so the min cut is the sink corner's two in-arcs: value = 2 * c2 * a * a.
    g, value = fixture_graph(mf, golden, "genrmf_a10_b1.max")
    assert value == 2 * 10 * 10 * 10
    g2 = mf.gen_genrmf(6, 1, 1, 7, seed=3)
    assert orc.dinic_maxflow(g2) == 2 * 7 * 36


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref",
                                                    "maxflow-serial")), reason="reference not built")
@pytest.mark.parametrize("name", ["test_DIMACS.txt", "ww.max", "m30.max", "snap_8.txt", "snap_4.txt"])
def test_reference_binary_matches(mf, orc, golden, name):
    rec = golden["fixtures"][name]
    ref, _ = orc.ref_serial(os.path.join(GOLDEN, name), rec["type"], rec["s"], rec["t"])
    assert ref == rec["value"]


def test_brute_force_min_cut(mf, orc):
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(3, 10))
        g = random_graph(mf, rng, n, int(rng.integers(n, 4 * n)), parallel=bool(trial % 2))
        bf = orc.brute_force_min_cut(g)
        assert orc.dinic_maxflow(g) == bf
        if trial % 2 == 0:  # the serial restatement aliases parallel arcs like the reference
            ser = orc.serial_maxflow(g)
            assert ser is None or ser == bf


def test_dinic_matches_scipy(mf, orc):
    sp = pytest.importorskip("scipy.sparse")
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    rng = np.random.default_rng(5)
    for _ in range(10):
        n = int(rng.integers(20, 200))
        g = random_graph(mf, rng, n, 6 * n, cap_max=1000)
        tails = g.tails()
        mat = sp.csr_matrix((g.cap.astype(np.int32), (tails, g.dst)), shape=(n, n))
        ref = csgraph.maximum_flow(mat, 0, n - 1, method="dinic").flow_value
        assert orc.dinic_maxflow(g) == ref


def test_serial_quirk_parallel_source_arcs(mf, orc):
    # preflow overwrites e(v) for parallel s->v arcs (maxflow-serial/graph.cpp:473): the
    # restatement keeps the reference's undercount, Dinic gives the true value.
    g = mf.Graph.from_arcs(3, [0, 0, 1], [1, 1, 2], [5, 7, 100], 0, 2)
    assert orc.dinic_maxflow(g) == 12
    assert orc.serial_maxflow(g) == 7


def test_no_path_returns_none(mf, orc):
    g = mf.Graph.from_arcs(4, [0, 2], [1, 3], [5, 5], 0, 3)
    assert orc.serial_maxflow(g) is None
    assert orc.dinic_maxflow(g) == 0


def test_certificate_rejects_planted_mistakes(mf, orc, golden):
    g, value = generated_graph(mf, golden, "genrmf8_s7")
    v, flow = orc.dinic_maxflow(g, with_flow=True)
    assert orc.certify(g, flow, v) == ("ok", value)
    bad = flow.copy()
    i = int(np.argmax(flow))
    bad[i] = g.cap[i] + 1
    assert orc.certify(g, bad, v)[0] == "capacity"
    bad = flow.copy()
    bad[i] -= 1
    assert orc.certify(g, bad, v)[0] in ("conservation", "value")
    assert orc.certify(g, flow, v - 1)[0] == "value"
    # a feasible but non-maximum flow (all zero) fails the cut check
    zero = np.zeros_like(flow)
    assert orc.certify(g, zero, 0)[0] == "cut"


def test_certificate_preflow_mode(mf, orc):
    # s -> a (5), a -> t (3), a -> b (4): a maximum preflow may leave 2 units at a or b.
    g = mf.Graph.from_arcs(4, [0, 1, 1], [1, 3, 2], [5, 3, 4], 0, 3)
    pre = np.array([5, 3, 0], dtype=np.int32)  # excess 2 stranded at a
    assert orc.certify(g, pre, 3, preflow=True)[0] == "ok"
    assert orc.certify(g, pre, 3, preflow=False)[0] == "conservation"


def test_rcsr_builder_restatement(mf, orc, golden):
    g, _ = generated_graph(mf, golden, "rlg16_s3")
    roff, rdst, fidx = orc.rcsr_build(g)
    assert roff[-1] == g.m
    tails = g.tails()
    # flow_index points at the arc rdst -> v for every reverse position of v
    v_of_k = np.repeat(np.arange(g.n), np.diff(roff))
    assert np.array_equal(g.dst[fidx], v_of_k)
    assert np.array_equal(tails[fidx], rdst)
