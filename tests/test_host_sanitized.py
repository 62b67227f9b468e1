"""The host loaders, writers, generators, terminal transforms and layout builders
(wbpr_amd/csrc/host_graph.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer: every
loader reads every fixture the reference's tests hold plus truncated, mutated and malformed
copies, and every graph that parses goes through the layout builders, the SNAP-binary and
DIMACS round trips and the terminal transforms.  Host code only (g++); no GPU."""
import os
import random
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def malformed_inputs(out):
    rng = random.Random(1)
    names = [f for f in os.listdir(GOLDEN) if f.endswith((".max", ".txt"))]
    for f in names:
        shutil.copy(os.path.join(GOLDEN, f), out)
        data = open(os.path.join(GOLDEN, f), "rb").read()
        for k in range(4):
            with open(os.path.join(out, f"trunc{k}_{f}"), "wb") as fh:
                fh.write(data[:rng.randint(0, len(data))])
        mutated = bytearray(data)
        for _ in range(20):
            if mutated:
                mutated[rng.randrange(len(mutated))] = rng.choice(b"0123456789 -\n\tpaecs%x")
        with open(os.path.join(out, f"mut_{f}"), "wb") as fh:
            fh.write(bytes(mutated))
    cases = {
        "empty.txt": b"",
        "hdr_huge.max": b"p max 2147483647 2147483647\nn 1 s\nn 2 t\na 1 2 5\n",
        "hdr_neg.max": b"p max -5 -3\nn 1 s\na 1 2 5\n",
        "bad_vertex.max": b"p max 3 2\nn 1 s\nn 3 t\na 1 7 5\na 0 2 3\n",
        "neg_cap.max": b"p max 3 2\nn 1 s\nn 3 t\na 1 2 -5\na 2 3 3\n",
        "snap_neg.txt": b"# c\n0 1\n-1 2\n3 4294967297\n",
        "konect.txt": b"% bip\n% 3 2 2\n1 1\n2 2\n3 1\n0 5\n",
        "garbage.bin": bytes(rng.randrange(256) for _ in range(4096)),
    }
    for name, data in cases.items():
        with open(os.path.join(out, name), "wb") as fh:
            fh.write(data)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_sanitizers(tmp_path):
    exe = tmp_path / "loaders_driver"
    build = subprocess.run(
        ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-pthread", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
         os.path.join(HERE, "host_sanitize", "loaders_driver.cpp"),
         os.path.join(ROOT, "wbpr_amd", "csrc", "host_graph.cpp"), "-o", str(exe)],
        capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-4000:]
    inputs = tmp_path / "inputs"
    scratch = tmp_path / "scratch"
    inputs.mkdir()
    scratch.mkdir()
    malformed_inputs(str(inputs))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), str(inputs), str(scratch)], capture_output=True, text=True,
                       timeout=600, env=env)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "runtime error" not in report and "AddressSanitizer" not in report, report[-4000:]
    calls, parsed = [int(x) for x in report.split("loader calls ")[1].split(", parsed ")]
    assert calls >= 4 * 60 and parsed >= 10
