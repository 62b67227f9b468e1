"""Host-side parts of libmf without a GPU: loaders, writers, generators, residual layouts, the
C ABI surface and the CLI's argument handling."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, fixture_graph, generated_graph


# ---- loaders ------------------------------------------------------------------------------
def test_dimacs_loader_small(mf):
    g = mf.load_dimacs(os.path.join(GOLDEN, "test_DIMACS.txt"))
    assert (g.n, g.m, g.s, g.t) == (4, 5, 0, 3)
    # CSR in vertex order, file order within a vertex (graph.cpp:200-208)
    assert g.offsets.tolist() == [0, 2, 3, 4, 5]
    assert g.dst.tolist() == [1, 2, 3, 3, 0]
    assert g.cap.tolist() == [10, 15, 20, 25, 5]
    assert g.m_header == 5


def test_snap_txt_loader(mf):
    g = mf.load_snap_txt(os.path.join(GOLDEN, "snap_4.txt"))
    # '# Nodes: 4 Edges: 6', unit capacities, sorted adjacency (graph.cpp:28-67)
    assert (g.n, g.m, g.s, g.t) == (4, 6, -1, -1)
    assert g.offsets.tolist() == [0, 2, 4, 6, 6]
    assert g.dst.tolist() == [1, 2, 2, 3, 1, 3]
    assert set(g.cap.tolist()) == {1}
    assert g.m_header == 6 and g.m_processed == 6


def test_snap_binary_roundtrip(mf, tmp_path):
    g = mf.load_snap_txt(os.path.join(GOLDEN, "snap_8.txt"))
    p = str(tmp_path / "g.bin")
    g.with_terminals(0, 7).save_snap_bin(p)
    raw = np.fromfile(p, dtype=np.int32)
    assert raw[:3].tolist() == [8, 11, 11]  # [n][m][m_processed]
    h = mf.load_snap_bin(p)
    assert h.n == g.n and np.array_equal(h.offsets, g.offsets) and np.array_equal(h.dst, g.dst)
    assert set(h.cap.tolist()) == {1}


def test_dimacs_roundtrip(mf, golden, tmp_path):
    g, _ = fixture_graph(mf, golden, "m10.max")
    p = str(tmp_path / "g.max")
    g.save_dimacs(p)
    h = mf.load_dimacs(p)
    assert (h.n, h.s, h.t) == (g.n, g.s, g.t)
    for a, b in ((h.offsets, g.offsets), (h.dst, g.dst), (h.cap, g.cap)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("text,status", [
    ("p max 3 1\nn 1 s\nn 3 t\na 1 4 5\n", "MF_ERR_PARSE"),     # vertex out of range
    ("p max 3 1\nn 1 s\nn 3 t\na 1 2 -5\n", "MF_ERR_PARSE"),    # negative capacity
    ("p max 3 1\na 1 2\n", "MF_ERR_PARSE"),                     # missing capacity
    ("c no problem line\na 1 2 3\n", "MF_ERR_PARSE"),           # no 'p' line
])
def test_dimacs_errors(mf, tmp_path, text, status):
    p = tmp_path / "bad.max"
    p.write_text(text)
    with pytest.raises(mf.MFError) as ei:
        mf.load_dimacs(str(p))
    assert ei.value.name == status


def test_missing_files(mf, tmp_path):
    for fn in (mf.load_dimacs, mf.load_snap_txt, mf.load_snap_bin):
        with pytest.raises(mf.MFError) as ei:
            fn(str(tmp_path / "nope"))
        assert ei.value.name == "MF_ERR_IO"


def test_snap_errors(mf, tmp_path):
    p = tmp_path / "a.txt"
    p.write_text("0 1\n1 2\n")  # no '# Nodes:' header
    with pytest.raises(mf.MFError):
        mf.load_snap_txt(str(p))
    p.write_text("# Nodes: 3 Edges: 1\n0 7\n")  # endpoint out of range
    with pytest.raises(mf.MFError):
        mf.load_snap_txt(str(p))
    b = tmp_path / "b.bin"
    np.array([4, 2, 2, 1, 9, 0, 1, 2, 2, 2], dtype=np.int32).tofile(str(b))  # dst 9 >= n
    with pytest.raises(mf.MFError):
        mf.load_snap_bin(str(b))
    # a header announcing more than the file holds fails before anything is allocated
    np.array([1 << 30, 1 << 30, 1 << 30, 0, 1], dtype=np.int32).tofile(str(b))
    with pytest.raises(mf.MFError) as ei:
        mf.load_snap_bin(str(b))
    assert ei.value.name == "MF_ERR_PARSE"
    g = mf.load_snap_txt(os.path.join(GOLDEN, "snap_8.txt"))
    g.with_terminals(0, 7).save_snap_bin(str(b))
    whole = b.read_bytes()
    b.write_bytes(whole[:-4])  # last offset missing
    with pytest.raises(mf.MFError) as ei:
        mf.load_snap_bin(str(b))
    assert ei.value.name == "MF_ERR_PARSE"


# ---- generators ---------------------------------------------------------------------------
def test_washington_rlg_structure(mf):
    g = mf.gen_washington_rlg(64, 64, 10000, seed=1)
    assert (g.n, g.m, g.s, g.t) == (4098, 12224, 0, 4097)  # 2*64 + 3*64*63 arcs
    tails = g.tails()
    src = g.cap[tails == 0]
    assert len(src) == 64 and set(src.tolist()) == {30000}  # 3r on source arcs
    mid = (tails > 0) & (g.dst != 4097)
    assert g.cap[mid].min() >= 1 and g.cap[mid].max() <= 10000
    for u in range(1, 64 * 63 + 1, 97):  # three distinct targets in the next level
        tgt = g.dst[g.offsets[u]:g.offsets[u + 1]]
        assert len(tgt) == 3 and len(set(tgt.tolist())) == 3
        lvl = (u - 1) // 64
        assert all((t - 1) // 64 == lvl + 1 for t in tgt)


def test_genrmf_structure(mf):
    a, b = 8, 5
    g = mf.gen_genrmf(a, b, 1, 100, seed=4)
    assert g.n == a * a * b and g.m == 5 * a * a * b - 4 * a * b - a * a
    assert (g.s, g.t) == (0, g.n - 1)
    big = g.cap == 100 * a * a
    assert big.sum() == 4 * a * (a - 1) * b  # in-frame 4-neighbour arcs
    inter = ~big
    assert inter.sum() == a * a * (b - 1)
    tails = g.tails()
    assert np.all(tails[inter] // (a * a) + 1 == g.dst[inter] // (a * a))
    # each frame -> frame link is a permutation
    for z in range(b - 1):
        sel = inter & (tails // (a * a) == z)
        assert len(set(g.dst[sel].tolist())) == a * a


def test_rmat_structure(mf):
    g = mf.gen_rmat(12, 16, seed=3)
    tails = g.tails()
    assert np.all(tails != g.dst)
    key = tails.astype(np.int64) * g.n + g.dst
    assert len(np.unique(key)) == g.m  # no duplicates
    assert g.cap.min() >= 1 and g.cap.max() <= 1000
    outdeg = np.diff(g.offsets)
    indeg = np.bincount(g.dst, minlength=g.n)
    assert outdeg[g.s] == outdeg.max()
    assert indeg[g.t] == np.delete(indeg, g.s).max()


def test_bipartite_structure(mf):
    L, R = 300, 200
    g = mf.gen_bipartite(L, R, 8.0, seed=2)
    assert g.n == L + R + 2 and (g.s, g.t) == (0, L + R + 1)
    assert set(g.cap.tolist()) == {1}
    tails = g.tails()
    lr = (tails >= 1) & (tails <= L)
    assert np.all((g.dst[lr] > L) & (g.dst[lr] <= L + R))
    assert abs(lr.sum() / L - 8.0) < 1.0
    for u in range(1, L + 1, 17):
        t = g.dst[g.offsets[u]:g.offsets[u + 1]]
        assert len(set(t.tolist())) == len(t)


def test_generators_deterministic(mf):
    a = mf.gen_genrmf(6, 4, 1, 50, seed=9)
    b = mf.gen_genrmf(6, 4, 1, 50, seed=9)
    c = mf.gen_genrmf(6, 4, 1, 50, seed=10)
    assert np.array_equal(a.cap, b.cap) and np.array_equal(a.dst, b.dst)
    assert not np.array_equal(a.cap, c.cap)


def test_generated_golden_values(mf, orc, golden):
    for name in ("rlg16_s3", "genrmf8_s7", "rmat10_s5", "bip_500_700_s5"):
        g, value = generated_graph(mf, golden, name)
        rec = golden["generated"][name]
        assert (g.n, g.m, g.s, g.t) == (rec["n"], rec["m"], rec["s"], rec["t"])
        assert orc.dinic_maxflow(g) == value


# ---- residual layouts -----------------------------------------------------------------------
@pytest.mark.parametrize("name", ["rlg16_s3", "genrmf8_s7", "rmat10_s5"])
def test_slot_layout_invariants(mf, golden, name):
    g, _ = generated_graph(mf, golden, name)
    L = mf.build_layout(g, mf.LAYOUT_BCSR)
    S = 2 * g.m
    off, dst, rev, res, fslot = L["offsets"], L["dst"], L["rev"], L["residual"], L["fslot"]
    assert off[-1] == S and len(dst) == S
    assert np.array_equal(rev[rev], np.arange(S))  # involution
    owner = np.repeat(np.arange(g.n), np.diff(off))
    assert np.array_equal(dst[rev], owner)  # reverse slot points back
    assert np.array_equal(res + res[rev], np.where(res > 0, res, res[rev]))
    tails = g.tails()
    assert np.array_equal(owner[fslot], tails) and np.array_equal(dst[fslot], g.dst)
    assert np.array_equal(res[fslot], g.cap) and np.all(res[rev[fslot]] == 0)


@pytest.mark.parametrize("name", ["rlg16_s3", "bip_500_700_s5", "rmat10_s5"])
def test_bcsr_matches_reference_builder(mf, orc, golden, name):
    # arcs in CSR order are the reference loader's file order for a tail-sorted file
    g, _ = generated_graph(mf, golden, name)
    L = mf.build_layout(g, mf.LAYOUT_BCSR)
    off, dst, res, bidx = orc.bcsr_build(g.n, g.tails(), g.dst, g.cap)
    assert np.array_equal(L["offsets"], off)
    assert np.array_equal(L["dst"], dst)
    assert np.array_equal(L["residual"], res)
    # the reference's linear search finds the paired slot when there are no antiparallel arcs
    tails = g.tails()
    key = set((tails.astype(np.int64) * g.n + g.dst).tolist())
    anti = any(int(v) * g.n + int(u) in key for u, v in zip(tails, g.dst))
    if not anti:
        assert np.array_equal(L["rev"], bidx)


@pytest.mark.parametrize("name", ["rlg16_s3", "genrmf8_s7", "rmat10_s5"])
def test_rcsr_matches_reference_builder(mf, orc, golden, name):
    g, _ = generated_graph(mf, golden, name)
    L = mf.build_layout(g, mf.LAYOUT_RCSR)
    roff, rdst, fidx = orc.rcsr_build(g)
    assert np.array_equal(L["roffsets"], roff)
    assert np.array_equal(L["rdst"], rdst)
    assert np.array_equal(L["flow_index"], fidx)  # no parallel arcs: identical


def test_layout_rejects_bad_graphs(mf):
    g = mf.Graph.from_arcs(3, [0, 1], [1, 2], [1, -1], 0, 2)
    with pytest.raises(mf.MFError):
        mf.build_layout(g, mf.LAYOUT_BCSR)
    g = mf.Graph.from_arcs(3, [0, 1], [1, 2], [1, 1], 1, 1)  # s == t
    with pytest.raises(mf.MFError):
        mf.build_layout(g, mf.LAYOUT_BCSR)


# ---- C ABI ----------------------------------------------------------------------------------
def test_every_declared_symbol_is_exported(mf):
    header = open(os.path.join(ROOT, "include", "mf.h")).read()
    declared = set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    missing = [s for s in sorted(declared) if not hasattr(mf._lib, s)]
    assert not missing, missing
    assert set(mf.EXPORTED) == declared


def test_status_strings_and_version(mf):
    assert "gfx950" in mf.version()
    for code in mf.STATUS:
        assert mf._lib.mf_status_string(code)
    assert mf.footprint_bytes(100, 1000) > 0


def test_no_gpu_context_error(mf):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(mf.MFError) as ei:
        mf.Context(0)
    assert ei.value.name in ("MF_ERR_NO_DEVICE", "MF_ERR_DEVICE")


# ---- CLI --------------------------------------------------------------------------------------
def run_cli(*args):
    return subprocess.run([os.path.join(ROOT, "wbpr_amd", "bin", "maxflow"), *args],
                          capture_output=True, text=True, timeout=120)


def test_cli_help():
    r = run_cli("-h")
    assert r.returncode == 0
    assert "-v\t\tSpecify which kind of file to load" in r.stdout
    assert "1: Vertex-centric push-relabel" in r.stdout


def test_cli_errors(tmp_path):
    assert run_cli("-v", "7", "-f", "x").returncode == 1          # invalid file type
    assert run_cli("-v", "2", "-f", str(tmp_path / "missing")).returncode == 1
    assert run_cli().returncode == 1                              # no input
    r = run_cli("-v", "2", "-f", os.path.join(GOLDEN, "test_DIMACS.txt"), "-a", "5")
    assert r.returncode == 1 and "Invalid algorithm type" in r.stdout
    r = run_cli("-v", "0", "-f", os.path.join(GOLDEN, "snap_4.txt"), "-s", "2", "-t", "2")
    assert r.returncode == 1


KONECT_SMALL = """% bip unweighted
% 9 4 3
1 1
1 2
2 1
3 2
3 3
4 3
4 3
2 2
4 1
"""


def test_konect_bipartite_loader(mf, orc, tmp_path):
    # the paper's bipartite-matching inputs: super-source -> L, L -> R, R -> super-sink
    import scipy.sparse as sp
    from scipy.sparse.csgraph import maximum_bipartite_matching
    p = tmp_path / "out.small"
    p.write_text(KONECT_SMALL)
    g = mf.load_konect_bipartite(str(p))
    L, R = 4, 3
    assert g.n == L + R + 2 and g.s == 0 and g.t == L + R + 1 and g.m_header == 9
    assert g.m == L + 9 + R and set(g.cap.tolist()) == {1}
    tails = g.tails()
    assert sorted(g.dst[tails == 0].tolist()) == [1, 2, 3, 4]
    assert np.all(g.dst[(tails >= L + 1) & (tails <= L + R)] == g.t)
    rows = [int(x.split()[0]) - 1 for x in KONECT_SMALL.splitlines() if x and x[0] != "%"]
    cols = [int(x.split()[1]) - 1 for x in KONECT_SMALL.splitlines() if x and x[0] != "%"]
    bi = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(L, R))
    matching = int((maximum_bipartite_matching(bi, perm_type="column") >= 0).sum())
    assert orc.dinic_maxflow(g) == matching == 3
    assert mf.load(str(p), 3).m == g.m


def test_konect_errors(mf, tmp_path):
    p = tmp_path / "bad"
    p.write_text("% bip\n1 x\n")
    with pytest.raises(mf.MFError) as ei:
        mf.load_konect_bipartite(str(p))
    assert ei.value.name == "MF_ERR_PARSE"
    with pytest.raises(mf.MFError):
        mf.load_konect_bipartite(str(tmp_path / "missing"))


def test_super_terminals(mf, orc):
    g = mf.gen_rmat(10, 8, seed=3)
    h = mf.super_terminals(g, [1, 2, 3], [7, 8])
    assert h.n == g.n + 2 and h.s == g.n and h.t == g.n + 1 and h.m == g.m + 5
    tails = h.tails()
    out_cap = np.bincount(g.tails(), weights=g.cap, minlength=g.n)
    in_cap = np.bincount(g.dst, weights=g.cap, minlength=g.n)
    sup = tails == h.s
    assert sorted(h.dst[sup].tolist()) == [1, 2, 3]
    for v, c in zip(h.dst[sup], h.cap[sup]):
        assert c == out_cap[v]
    for u in (7, 8):
        sel = (tails == u) & (h.dst == h.t)
        assert sel.sum() == 1 and h.cap[sel][0] == in_cap[u]
    # the super arcs never bind: the value is the max flow between the terminal sets
    assert orc.dinic_maxflow(h) <= min(out_cap[[1, 2, 3]].sum(), in_cap[[7, 8]].sum())
    with pytest.raises(mf.MFError):
        mf.super_terminals(g, [1, 1], [7])  # a vertex used twice
    with pytest.raises(mf.MFError):
        mf.super_terminals(g, [1], [g.n])  # out of range


def test_bfs_terminal_pairs(mf):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    g = mf.gen_rmat(11, 8, seed=4)
    h, src, snk = mf.bfs_terminal_pairs(g, 20, seed=9)
    h2, src2, snk2 = mf.bfs_terminal_pairs(g, 20, seed=9)
    assert np.array_equal(src, src2) and np.array_equal(snk, snk2) and h.m == h2.m
    assert len(set(src.tolist()) | set(snk.tolist())) == 40  # disjoint pairs
    assert h.n == g.n + 2 and h.s == g.n and h.t == g.n + 1
    A = csr_matrix((np.ones(g.m), (g.tails(), g.dst)), shape=(g.n, g.n))
    d = shortest_path(A, unweighted=True, indices=src)
    for i, (s, t) in enumerate(zip(src, snk)):
        di = d[i][np.isfinite(d[i])]
        dmax = int(di.max())
        assert np.isfinite(d[i][t]) and d[i][t] >= max(dmax - dmax // 4, 1)
    empty = mf.Graph.from_arcs(10, np.array([0]), np.array([1]), np.array([1]), 0, 1)
    with pytest.raises(mf.MFError) as ei:
        mf.bfs_terminal_pairs(empty, 3, seed=1)
    assert ei.value.name == "MF_ERR_NO_PATH"
