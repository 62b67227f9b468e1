"""Writes tests/golden/values.json: exact max-flow values of the fixture files and of the
seeded generator instances the tests and the benchmark use.

Every value is computed by the CPU oracle's Dinic and, where it finishes quickly, checked
against the C restatement of the reference serial solver and against the reference's own
serial binary (oracle/_ref/maxflow-serial, built from the reference sources).  Disagreement
aborts.  Run from the repository root:  python tests/golden/make_golden.py [--big]
(--big adds the large benchmark instances; Genrmf 64x64 takes about a minute of Dinic,
Genrmf 128x128 about half an hour).
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import wbpr_amd as mf  # noqa: E402
from oracle import oracle as orc  # noqa: E402

# fixture files copied from the reference tree (data only): name -> (file type, s, t)
FIXTURES = {
    "test_DIMACS.txt": (2, None, None),   # maxflow-serial/input/test_DIMACS.txt
    "genrmf_a10_b1.max": (2, None, None),  # network/DIMACS/genrmf/test_graph.txt
    "genrmf_a4_b3.max": (2, None, None),   # network/DIMACS/genrmf/network.dimacs
    "ww.max": (2, None, None),             # solver/solver-4/ww.max
    "m10.max": (2, None, None),            # solver/solver-3/m10.max
    "m30.max": (2, None, None),            # solver/solver-3/m30.max
    "m50.max": (2, None, None),            # solver/solver-3/m50.max
    "waissi_example.max": (2, None, None),  # network/DIMACS/waissi/tp55/example.max
    "snap_8.txt": (0, 0, 7),               # maxflow-cuda/test/test_graph.txt
    "snap_5.txt": (0, 0, 4),               # maxflow-serial/test_graph.txt
    "snap_4.txt": (0, 0, 3),               # maxflow-cuda/snap_test.txt
}

# generated instances: name -> (generator, args)
GENERATED = {
    "rlg16_s3": ("gen_washington_rlg", (16, 16, 10000, 3)),
    "rlg32_s7": ("gen_washington_rlg", (32, 16, 10000, 7)),
    "rlg64_s1": ("gen_washington_rlg", (64, 64, 10000, 1)),
    "genrmf8_s7": ("gen_genrmf", (8, 8, 1, 10000, 7)),
    "genrmf12_s2": ("gen_genrmf", (12, 6, 1, 10000, 2)),
    "genrmf14_s1": ("gen_genrmf", (14, 14, 1, 10000, 1)),
    "genrmf16_s1": ("gen_genrmf", (16, 16, 1, 10000, 1)),
    "rmat10_s5": ("gen_rmat", (10, 16, 0.57, 0.19, 0.19, 1000, 5)),
    "rmat12_s3": ("gen_rmat", (12, 16, 0.57, 0.19, 0.19, 1000, 3)),
    "rmat16_s1": ("gen_rmat", (16, 16, 0.57, 0.19, 0.19, 1000, 1)),
    "bip_500_700_s5": ("gen_bipartite", (500, 700, 8.0, 5)),
    "bip_3000_2000_s9": ("gen_bipartite", (3000, 2000, 8.0, 9)),
}
BIG = {
    "genrmf64_s1": ("gen_genrmf", (64, 64, 1, 10000, 1)),
    "rmat22_s1": ("gen_rmat", (22, 16, 0.57, 0.19, 0.19, 1000, 1)),
    "genrmf128_s1": ("gen_genrmf", (128, 128, 1, 10000, 1)),  # about 30 minutes of Dinic
}
SERIAL_MAX_N = 3000  # the serial solvers are O(n) per operation


def value_of(g, name):
    v = orc.dinic_maxflow(g)
    if g.n <= SERIAL_MAX_N:
        ser = orc.serial_maxflow(g)
        assert ser == v, f"{name}: serial restatement {ser} != dinic {v}"
        if orc.ref_available():
            with tempfile.TemporaryDirectory() as tmp:
                p = os.path.join(tmp, "g.max")
                g.save_dimacs(p)
                ref, _ = orc.ref_serial(p, 2, timeout=600)
            assert ref == v, f"{name}: reference binary {ref} != dinic {v}"
    return v


def main():
    out = {"fixtures": {}, "generated": {}}
    for name, (ft, s, t) in FIXTURES.items():
        path = os.path.join(HERE, name)
        g = mf.load(path, ft)
        if ft != 2:
            g = g.with_terminals(s, t)
        v = value_of(g, name)
        if orc.ref_available():
            ref, _ = orc.ref_serial(path, ft, g.s, g.t)
            assert ref == v, f"{name}: reference binary {ref} != {v}"
        out["fixtures"][name] = {"type": ft, "s": g.s, "t": g.t, "n": g.n, "m": g.m, "value": v}
        print(name, out["fixtures"][name], flush=True)
    gens = dict(GENERATED)
    if "--big" in sys.argv:
        gens.update(BIG)
    for name, (fn, args) in gens.items():
        g = getattr(mf, fn)(*args)
        v = value_of(g, name)
        out["generated"][name] = {"gen": fn, "args": list(args), "n": g.n, "m": g.m,
                                  "s": g.s, "t": g.t, "value": v}
        print(name, out["generated"][name], flush=True)
    path = os.path.join(HERE, "values.json")
    old = {}
    if os.path.exists(path):
        with open(path) as f:
            old = json.load(f)
    for k in ("generated",):
        for name, rec in old.get(k, {}).items():
            out[k].setdefault(name, rec)  # keep big entries computed earlier
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
