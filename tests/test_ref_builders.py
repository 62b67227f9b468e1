"""The library's loaders and residual layouts against the reference's own host code.

oracle/Makefile compiles the reference's unmodified src/graph.cpp of maxflow-cuda (RCSR),
maxflow-bcsr (BCSR) and maxflow-bcsr2 (BCSR2) with a small dump driver (oracle/ref_layouts.cpp)
into oracle/_ref/.  These tests feed the same files to both sides and compare arrays:

  loaders   CSRGraph::buildFromDIMACSFile / buildFromTxtFile / loadFromBinary
            (maxflow-cuda/src/graph.cpp:28-67,151-214,292-314) against mf.load_*;
  RCSR      ResidualGraph::buildFromCSRGraph (maxflow-cuda/src/graph.cpp:317-402):
            roffsets, rdestinations, flow_index against mf.build_layout(RCSR);
  BCSR      the BCSR DIMACS loader, which appends a reverse slot per arc
            (maxflow-bcsr/src/graph.cpp:151-219), and its backward_idx search (:322-365)
            against mf.build_layout(BCSR) offsets / dst / residual / rev;
  BCSR2     the interleaved {residual, reverse slot} pairs (maxflow-bcsr2/src/graph.cpp:322-365).

The reference pairs a slot with the first slot of the head's list that points back
(graph.cpp:351-362 / :386-399): with parallel or antiparallel arcs that is not the arc's own
reverse slot (it aliases them), so the reverse-slot arrays are compared on graphs without such
arcs; everything else is compared on every graph.  Layout files are written tail-sorted (the
order mf.save_dimacs writes), which is the order in which both sides lay out slots.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, fixture_graph, generated_graph

pytestmark = pytest.mark.skipif(
    not __import__("oracle.oracle", fromlist=["x"]).ref_layouts_available(),
    reason="reference sources were not present when oracle/ was built (oracle/_ref/ref_layouts_*)")

DIMACS_FIXTURES = ["genrmf_a10_b1.max", "genrmf_a4_b3.max", "m10.max", "m30.max", "m50.max",
                   "test_DIMACS.txt", "waissi_example.max", "ww.max"]
SNAP_FIXTURES = ["snap_4.txt", "snap_5.txt", "snap_8.txt"]
LAYOUT_GRAPHS = ["rlg16_s3", "rlg32_s7", "bip_500_700_s5", "bip_3000_2000_s9", "rmat10_s5",
                 "genrmf8_s7", "genrmf12_s2"]


def arc_flags(g):
    """(has parallel arcs, has antiparallel arcs)."""
    tails = g.tails().astype(np.int64)
    key = tails * g.n + g.dst
    par = len(np.unique(key)) != len(key)
    rkey = g.dst.astype(np.int64) * g.n + tails
    anti = bool(np.isin(rkey, key).any())
    return par, anti


@pytest.mark.parametrize("name", DIMACS_FIXTURES)
def test_dimacs_loader_matches_reference(mf, orc, name, tmp_path):
    path = os.path.join(GOLDEN, name)
    ref = orc.ref_layouts("rcsr", path, 2, str(tmp_path))
    g = mf.load_dimacs(path)
    n, m, _, s, t = ref["csr_header"].tolist()
    assert (g.n, g.m, g.s, g.t) == (n, m, s, t)
    assert np.array_equal(g.offsets, ref["csr_offsets"])
    assert np.array_equal(g.dst, ref["csr_destinations"])
    assert np.array_equal(g.cap, ref["csr_capacities"])


@pytest.mark.parametrize("name", SNAP_FIXTURES)
def test_snap_txt_loader_matches_reference(mf, orc, name, tmp_path):
    path = os.path.join(GOLDEN, name)
    ref = orc.ref_layouts("rcsr", path, 0, str(tmp_path))
    g = mf.load_snap_txt(path)
    n, m, m_processed, _, _ = ref["csr_header"].tolist()
    assert (g.n, g.m_header, g.m_processed) == (n, m, m_processed)
    assert np.array_equal(g.offsets, ref["csr_offsets"])
    assert np.array_equal(g.dst, ref["csr_destinations"])
    assert np.array_equal(g.cap, ref["csr_capacities"])


@pytest.mark.parametrize("name", SNAP_FIXTURES + ["rmat10_s5"])
def test_snap_binary_matches_reference(mf, orc, golden, name, tmp_path):
    if name.endswith(".txt"):
        g = mf.load_snap_txt(os.path.join(GOLDEN, name))
    else:
        g, _ = generated_graph(mf, golden, name)
    path = str(tmp_path / "g.bin")
    g.save_snap_bin(path)
    ref = orc.ref_layouts("rcsr", path, 1, str(tmp_path))
    h = mf.load_snap_bin(path)
    n, m, m_processed, _, _ = ref["csr_header"].tolist()
    assert (h.n, h.m_header, h.m_processed) == (n, m, m_processed)
    assert np.array_equal(h.offsets, ref["csr_offsets"])
    assert np.array_equal(h.dst, ref["csr_destinations"])
    assert np.array_equal(h.cap, ref["csr_capacities"])  # unit capacities


def layout_graph(mf, golden, name):
    if name in golden["generated"]:
        return generated_graph(mf, golden, name)[0]
    return fixture_graph(mf, golden, name)[0]


@pytest.mark.parametrize("name", LAYOUT_GRAPHS + ["m30.max", "ww.max"])
def test_rcsr_layout_matches_reference(mf, orc, golden, name, tmp_path):
    g = layout_graph(mf, golden, name)
    path = str(tmp_path / "g.max")
    g.save_dimacs(path)
    ref = orc.ref_layouts("rcsr", path, 2, str(tmp_path))
    L = mf.build_layout(g, mf.LAYOUT_RCSR)
    assert np.array_equal(ref["offsets"], g.offsets)
    assert np.array_equal(ref["destinations"], g.dst)
    assert np.array_equal(ref["forward_flows"], g.cap)
    assert not ref["backward_flows"].any()
    assert np.array_equal(L["roffsets"], ref["roffsets"])
    assert np.array_equal(L["rdst"], ref["rdestinations"])
    par, _ = arc_flags(g)
    if not par:  # the reference's flow_index aliases parallel arcs to the first one
        assert np.array_equal(L["flow_index"], ref["flow_index"])


@pytest.mark.parametrize("name", LAYOUT_GRAPHS + ["m30.max", "ww.max"])
def test_bcsr_layouts_match_reference(mf, orc, golden, name, tmp_path):
    g = layout_graph(mf, golden, name)
    path = str(tmp_path / "g.max")
    g.save_dimacs(path)
    par, anti = arc_flags(g)
    L = mf.build_layout(g, mf.LAYOUT_BCSR)
    L2 = mf.build_layout(g, mf.LAYOUT_BCSR2)
    ref = orc.ref_layouts("bcsr", path, 2, str(tmp_path))
    ref2 = orc.ref_layouts("bcsr2", path, 2, str(tmp_path))
    for lay, r in ((L, ref), (L2, ref2)):
        assert np.array_equal(lay["offsets"], r["csr_offsets"])
        assert np.array_equal(lay["dst"], r["csr_destinations"])
        assert np.array_equal(lay["residual"], r["csr_capacities"])
        assert np.array_equal(r["offsets"], r["csr_offsets"])
    pairs = ref2["forward_flows"].reshape(-1, 2)
    assert np.array_equal(L2["residual"], pairs[:, 0])
    if not (par or anti):
        assert np.array_equal(L["rev"], ref["backward_idx"])
        assert np.array_equal(L2["rev"], pairs[:, 1])
    else:
        # the reference's pairing differs only where its search stops at another arc's slot
        assert (L["rev"] == ref["backward_idx"]).mean() > 0.5
