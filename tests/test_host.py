"""Host-side product checks that need no GPU: the C ABI library loads and exports every symbol
include/nq.h declares, the product generator equals the oracle/reference task lists, the C++
drop-in links, and the CLI reproduces the reference's argument validation."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import n_queens_amd as nq
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header: str) -> set[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nq_\w+)\s*\(", text, flags=re.M))


def test_library_exports_every_declared_symbol():
    L = nq.lib()
    decl = declared_functions("nq.h")
    assert decl, "no declarations parsed"
    assert decl == set(nq.EXPORTS), decl ^ set(nq.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", nq.LIB_PATH], capture_output=True, text=True).stdout
    # Every function of the reference header (src/n_queens.h:7-19) and of src/utils.h:5-7, with
    # the reference's C++ signatures.
    for cxx in ("_Z13cuda_n_queensii", "_Z16partial_n_queensiiiiRSt6vectorIiSaIiEEi",
                "_Z24partial_n_queens_for_oddiiiiRSt6vectorIiSaIiEEi", "_Z8n_queensiiiiRx",
                "_Z18n_queens_iterativeiiiiRx", "_Z15serial_n_queensii", "_Z17parallel_n_queensii",
                "_Z12time_diff_msR7timevalS0_", "_Z15print_with_timePKcz"):
        assert cxx in out, cxx


@pytest.mark.parametrize("N", range(2, 18))
def test_product_generator_equals_oracle(N):
    for rows in range(1, min(N, 7) + 1):
        assert np.array_equal(nq.generate(N, rows), oracle.gen_tasks(N, rows)), (N, rows)


def test_product_generator_vs_reference_vectors(ref_vectors):
    for N, rows, _ in ref_vectors["totals"]:
        N, rows = int(N), int(rows)
        assert np.array_equal(nq.generate(N, rows), ref_vectors[f"N{N}_r{rows}_tasks"])


def test_product_generator_task_counts_vs_logs(published):
    for N, rows, k, _src in published["task_counts"]:
        if N <= 22:
            assert len(nq.generate(N, rows)) == k, (N, rows)


def test_product_generator_large_matches_oracle():
    a = nq.generate(21, 6)
    assert len(a) == 4470757
    assert np.array_equal(a, oracle.gen_tasks(21, 6))


def test_generate_rejects_bad_args():
    with pytest.raises(nq.NQError):
        nq.generate(0, 1)
    with pytest.raises(nq.NQError):
        nq.generate(8, 9)


def test_cpp_dropin_links_and_generates(tmp_path):
    """A reference-style caller links against libnq_hip.so through include/n_queens.h."""
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "n_queens.h"\n#include <cstdio>\n'
        "int main(){ vector<int> t; partial_n_queens(9,0,0,0,t,3); partial_n_queens_for_odd(9,0,0,0,t,3);\n"
        ' printf("%zu", t.size()/3); for (int x : t) printf(" %d", x); printf("\\n"); return 0; }\n')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(nq.LIB_PATH)
    subprocess.check_call(["g++", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L", libdir, "-lnq_hip", f"-Wl,-rpath,{libdir}"])
    words = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    expect = oracle.gen_tasks(9, 3)
    assert words[0] == len(expect)
    assert np.array_equal(np.array(words[1:], dtype=np.int64).astype(np.uint32).reshape(-1, 3), expect)


def test_cpp_dropin_cpu_counters(tmp_path, published):
    """The reference header's host counters (src/n_queens.h:7-17): n_queens accumulates into the
    caller's sum; serial_n_queens and parallel_n_queens fold x2 over the generator's tasks."""
    src = tmp_path / "cpu.cpp"
    src.write_text(
        '#include "n_queens.h"\n#include <cstdio>\n'
        "int main(){ long long s = 5; n_queens(8, 0, 0, 0, s);\n"
        " vector<int> t; partial_n_queens(11,0,0,0,t,3); partial_n_queens_for_odd(11,0,0,0,t,3); long long p = 0;\n"
        " for (size_t i = 0; i < t.size(); i += 3) n_queens(11, t[i], t[i+1], t[i+2], p);\n"
        ' printf("%lld %lld", s, p);\n'
        " for (int N = 4; N <= 13; ++N) printf(\" %lld %lld\", serial_n_queens(N, 2), parallel_n_queens(N, N < 8 ? 2 : 4));\n"
        ' printf("\\n"); return 0; }\n')
    exe = tmp_path / "cpu"
    libdir = os.path.dirname(nq.LIB_PATH)
    subprocess.check_call(["g++", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L", libdir, "-lnq_hip", f"-Wl,-rpath,{libdir}"])
    w = [int(x) for x in subprocess.check_output([str(exe)], env=NO_GPU).split()]
    assert w[0] == 5 + 92                                   # accumulates into sum
    assert w[1] == oracle.count_tasks(11, oracle.gen_tasks(11, 3))
    for i, N in enumerate(range(4, 14)):
        assert w[2 + 2 * i] == w[3 + 2 * i] == published["solutions"][N], N


NO_GPU = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")


def test_cpp_dropin_iterative_counter(tmp_path):
    """n_queens_iterative (src/n_queens.cpp:24-72) against the oracle's restatement, on seeded
    prefixes of every depth: it accumulates into sum, counts the completions below N-1 queens
    and adds nothing for N-1 queens or a full board, as the reference's loop does."""
    rng = np.random.default_rng(0x17E6)
    cases = []
    while len(cases) < 200:
        N = int(rng.integers(4, 14))
        k = int(rng.integers(0, N + 1))
        cols = []
        for _ in range(k):
            free = [c for c in range(N) if oracle.encode_prefix(N, [*cols, c]) is not None]
            if not free:
                break
            cols.append(int(free[rng.integers(len(free))]))
        if len(cols) == k:
            cases.append((N, *oracle.encode_prefix(N, cols)))
    src = tmp_path / "iter.cpp"
    src.write_text(
        '#include "n_queens.h"\n#include <cstdio>\n#include <cstdlib>\n'
        "int main(int argc, char **argv){ for (int i = 1; i + 3 < argc; i += 4) { long long s = 7;\n"
        " n_queens_iterative(atoi(argv[i]), (int)strtoul(argv[i+1],0,10), (int)strtoul(argv[i+2],0,10),\n"
        ' (int)strtoul(argv[i+3],0,10), s); printf("%lld\\n", s - 7); } return 0; }\n')
    exe = tmp_path / "iter"
    libdir = os.path.dirname(nq.LIB_PATH)
    subprocess.check_call(["g++", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L", libdir, "-lnq_hip", f"-Wl,-rpath,{libdir}"])
    args = [str(x) for case in cases for x in case]
    got = [int(x) for x in subprocess.check_output([str(exe), *args], env=NO_GPU).split()]
    assert got == [oracle.count_iterative(*case) for case in cases]
    assert any(g == 0 and bin(c).count("1") >= N - 1 for g, (N, c, _, _) in zip(got, cases))


REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_SRC, "main.cpp")), reason="no reference checkout")
def test_reference_main_links_against_the_library(tmp_path):
    """The reference's own src/main.cpp, compiled as its compile.sh does (CONFIG2), links against
    libnq_hip.so alone (cuda_n_queens, print_with_time, time_diff_ms) and, with the device
    hidden, prints the reference's lines: the task count of (13, 5) and result -1."""
    exe = tmp_path / "ref_main"
    libdir = os.path.dirname(nq.LIB_PATH)
    subprocess.check_call(["g++", "-O2", "-D_USE_CONFIG2_", "-I", os.path.join(ROOT, "include"),
                           os.path.join(REF_SRC, "main.cpp"), "-o", str(exe), "-L", libdir, "-lnq_hip",
                           f"-Wl,-rpath,{libdir}"])
    p = subprocess.run([str(exe), "13", "5"], capture_output=True, text=True, env=NO_GPU, timeout=60)
    assert p.returncode == 0, p.stderr
    stamp = r"\[\d{4}-\d\d-\d\d \d\d:\d\d:\d\d\.\d{3}\] "
    assert re.search(stamp + r"Use [\d.]+ms to generate 15550 subproblems!", p.stdout), p.stdout
    assert "Failed to find any gpu!" in p.stdout
    assert re.search(stamp + r"cuda 13 queens result -1, calc time: \[[\d.]+ms\]", p.stdout), p.stdout
    assert len(oracle.gen_tasks(13, 5)) == 15550


def run_cli(*args, env=NO_GPU):
    # The device is hidden so that no case here starts a real count, also on a GPU machine.
    return subprocess.run([nq.CLI_PATH, *map(str, args)], capture_output=True, text=True, env=env, timeout=60)


def test_cli_validation_order():
    # src/main.cpp:10-29, in order.
    p = run_cli(29, 5)
    assert p.returncode == 255 and p.stdout.strip() == "Not allowed 29 queens problem!"
    p = run_cli(1, 0)
    assert p.stdout.strip() == "Not allowed 1 queens problem!"
    p = run_cli(25, 3)  # CONFIG2: STACKSIZE 19, 19+1+3 < 25
    assert p.stdout.strip() == ("This configure cannot process 25-queens with pre-placing 3 rows. "
                                "Pre-placed rows should be >= 5.")
    p = run_cli(25, 3, "--config", 1)  # CONFIG1: STACKSIZE 24 accepts it; rows fine -> would run
    assert "This configure" not in p.stdout
    p = run_cli(13, 12)
    assert p.stdout.strip() == "Pre-placed rows should be <= 11."


def test_cli_cpu_cross_checks():
    """--serial / --openmp print the reference's commented-out cross-check lines
    (src/main.cpp:35-47) before the GPU part."""
    p = run_cli(12, 4, "--serial", "--openmp")
    lines = p.stdout.splitlines()
    assert re.match(r"serial 12 queens result 14200, calc time: \[[\d.]+ms\]$", lines[0])
    assert re.match(r"parallel 12 queens result 14200, calc time: \[[\d.]+ms\]$", lines[1])
    assert "Failed to find any gpu!" in p.stdout


def test_cli_no_gpu_path():
    p = run_cli(13, 1)
    lines = p.stdout.splitlines()
    assert lines[0] == "Pre-placed rows shold be at least 2, so set to 2!"
    assert re.match(r"\[\d{4}-\d\d-\d\d \d\d:\d\d:\d\d\.\d{3}\] =+$", lines[1])
    assert re.search(r"Use [\d.]+ms to generate 66 subproblems!", lines[2])
    assert "Failed to find any gpu!" in p.stdout
    assert re.search(r"cuda 13 queens result -1, calc time: \[[\d.]+ms\]", p.stdout)


def test_tuner_candidates_respect_depth():
    """The autotuner only proposes splits that leave the DFS two rows to place."""
    from n_queens_amd import tune
    for N, rows in ((21, 6), (12, 9), (8, 5), (6, 5)):
        cands = tune.candidates(N, rows)
        assert cands
        assert all(rows + c.split + 2 < N or c.split == 0 for c in cands)
        assert {c.push for c in cands} == {nq.NQ_PUSH_MASKED, nq.NQ_PUSH_ALWAYS}
    assert len(tune.candidates(21, 6, quick=True)) == 8


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_resumable_rejects_foreign_checkpoints(tmp_path):
    """A checkpoint for another problem, another task count, or not a checkpoint at all is
    refused before any device work (so this runs without a GPU)."""
    p = str(tmp_path / "ck")
    _write(p, "nq-checkpoint 1\nN 13 rows 5 ntasks 15550 next 0 sum 0\n")
    with pytest.raises(nq.NQError) as e:
        nq.solve_resumable(14, 5, p)
    assert e.value.status == nq.NQ_ERR_ARG and "different problem" in str(e.value)
    _write(p, "nq-checkpoint 1\nN 13 rows 5 ntasks 999 next 0 sum 0\n")
    with pytest.raises(nq.NQError) as e:
        nq.solve_resumable(13, 5, p)
    assert "task count" in str(e.value)
    _write(p, "nq-checkpoint 1\nN 13 rows 5 ntasks 15550 next 20000 sum 0\n")
    with pytest.raises(nq.NQError) as e:
        nq.solve_resumable(13, 5, p)
    assert "past its task count" in str(e.value)
    _write(p, "hello\n")
    with pytest.raises(nq.NQError) as e:
        nq.solve_resumable(13, 5, p)
    assert "not a checkpoint" in str(e.value)


def test_resumable_writes_initial_checkpoint_without_gpu(tmp_path):
    """With no device the call fails with NQ_ERR_NO_DEVICE after writing the initial checkpoint
    (cursor 0, the generator's task count)."""
    p = str(tmp_path / "ck")
    r = subprocess.run([sys.executable, "-c", f"""
import n_queens_amd as nq
try:
    nq.solve_resumable(13, 5, {p!r}, range_tasks=1000)
except nq.NQError as e:
    print("status", e.status)
"""], capture_output=True, text=True, env=NO_GPU, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
        timeout=120)
    assert f"status {nq.NQ_ERR_NO_DEVICE}" in r.stdout, r.stdout + r.stderr
    assert open(p).read() == "nq-checkpoint 1\nN 13 rows 5 ntasks 15550 next 0 sum 0\n"


def test_tuner_occupancy_candidates():
    from n_queens_amd import tune
    win = nq.Config(block=128, chunk=16, split=2)
    occ = tune.occupancy_candidates(20, 6, win)
    assert [c.waves_per_cu for c in occ[:4]] == [4, 8, 12, 16]
    assert [c.stack_depth for c in occ[4:]] == [13, 15, 17]
    assert all(c.split == 2 and c.block == 128 for c in occ)


def test_default_and_solve_configs_leave_block_automatic():
    """nq_config_default / nq_solve_config need no device: workgroup size 0 (automatic, chosen by
    occupancy at launch), the smallest legal stack, chunk 16, automatic push, and the split rule
    clamp(N - rows - R, 0, 3) with R = (N+4)/2 up to N=20 and 13 from N=21 on."""
    c = nq.config_default(21, 9)
    assert (c.block, c.stack_depth, c.chunk, c.push, c.split) == (0, 11, 16, nq.NQ_PUSH_AUTO, 0)
    for N, rows, split in ((21, 6, 2), (22, 6, 3), (20, 6, 2), (19, 6, 2), (18, 6, 1), (17, 5, 2), (16, 4, 2),
                           (13, 5, 0), (28, 7, 3), (24, 10, 1)):
        s = nq.solve_config(N, rows)
        assert s.split == split and s.block == 0 and s.stack_depth == N - 1 - rows, (N, rows)


def test_config_resolve_needs_a_device():
    r = subprocess.run([sys.executable, "-c", """
import n_queens_amd as nq
try:
    nq.resolve_config(21, 9, 1 << 30)
except nq.NQError as e:
    print("status", e.status)
"""], capture_output=True, text=True, env=NO_GPU, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
        timeout=120)
    assert f"status {nq.NQ_ERR_NO_DEVICE}" in r.stdout, r.stdout + r.stderr
