"""The server's host side on the CPU: n_queens_amd/csrc/nq_serve.cpp (with the library's request
expansion) built against a mock HIP runtime whose kernel emulates serve_kernel's protocol
(tests/serve_model.cpp).  A small ring, few slots and a small task budget make the ring wrap,
slots get reused and submissions be refused as busy; three phases separated by idle gaps make the
kernel instance stop and restart on the other tail.  Every served count is checked against the
oracle, and a ThreadSanitizer build of the same model must report no race."""
import json
import os
import shutil
import subprocess

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "n_queens_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "serve_model.cpp")] + [
    os.path.join(CSRC, f) for f in ("nq_serve.cpp", "nq_batch.cpp", "nq_prefix.cpp")]
FLAGS = ["-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
         "-I" + CSRC, "-pthread"]
CLANG = "/opt/rocm/llvm/bin/clang++"


def build(tmp_path, compiler, extra, name):
    exe = tmp_path / name
    subprocess.check_call([compiler, *FLAGS, *extra, *SOURCES, "-o", str(exe)])
    return exe


def expected(N, cols, published):
    return published["solutions"][N] if not cols else oracle.count_subtree(N, *oracle.encode_prefix(N, cols))


def check_output(text, published, nreq):
    lines = text.strip().splitlines()
    assert len(lines) == nreq
    for line in lines:
        lhs, rhs = line.split("=")
        a = [int(x) for x in lhs.split()]
        N, k, cols = a[0], a[1], a[2:]
        assert len(cols) == k
        assert int(rhs) == expected(N, cols, published), line


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_served_counts_exact_through_wraps_slot_reuse_and_restarts(tmp_path, published, seed):
    exe = build(tmp_path, "g++", ["-O2"], "serve_model")
    trace = tmp_path / "trace.txt"
    p = subprocess.run([str(exe), str(seed), "300"], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, NQ_SERVE_TRACE=str(trace)))
    assert p.returncode == 0, p.stderr
    check_output(p.stdout, published, 300)
    kinds = [line.split()[1] for line in trace.read_text().splitlines()]
    # Three phases: at least three instances launched and stopped; many more slices than the
    # 6 slots (reuse), and the 512-entry ring wrapped (its last position is far past 512).
    assert kinds.count("L") >= 3 and kinds.count("S") >= 3, p.stderr
    assert kinds.count("P") == kinds.count("C") > 6 * 20
    last_end = max(int(line.split()[3]) for line in trace.read_text().splitlines() if line.split()[1] == "P")
    assert last_end > 512 * 10


@pytest.mark.skipif(not shutil.which(CLANG), reason="no clang++ with ThreadSanitizer")
def test_scheduler_has_no_data_race(tmp_path, published):
    """Host-only ThreadSanitizer build (the mock kernel's flush orders its ring reads before the
    completion, as the kernel's data dependency does)."""
    exe = build(tmp_path, CLANG, ["-O1", "-g", "-fsanitize=thread"], "serve_model_tsan")
    p = subprocess.run([str(exe), "9", "150"], capture_output=True, text=True, timeout=300)
    assert "ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
    assert p.returncode == 0, p.stderr[-2000:]
    check_output(p.stdout, published, 150)
