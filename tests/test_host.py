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


def test_generate_device_checks_args_before_the_device():
    """nq_generate_device validates N and rows as nq_generate does, and without a device
    reports NQ_ERR_NO_DEVICE."""
    for N, rows in ((0, 1), (32, 2), (8, 0), (8, 9)):
        with pytest.raises(nq.NQError) as e:
            nq.generate_device(N, rows)
        assert e.value.status == nq.NQ_ERR_ARG, (N, rows)
    r = subprocess.run([sys.executable, "-c", """
import n_queens_amd as nq
try:
    nq.generate_device(21, 6)
except nq.NQError as e:
    print("status", e.status)
"""], capture_output=True, text=True, env=NO_GPU, cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
        timeout=120)
    assert f"status {nq.NQ_ERR_NO_DEVICE}" in r.stdout, r.stdout + r.stderr


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


def test_tuner_search_space():
    """The autotuner's phases: preplace 5..7 x split 0..4 keeping three rows for the DFS, then
    workgroup size x push form, chunk 1..64, waves per CU 4..16 and extra LDS frames."""
    from n_queens_amd import tune
    for N in (21, 20, 12, 9):
        pts = tune.depth_points(N)
        assert pts and all(2 <= p.preplace and p.preplace + p.split <= N - 3 for p in pts), N
    assert {(p.preplace, p.split) for p in tune.depth_points(21)} == {(g, s) for g in (5, 6, 7) for s in range(5)}
    p = tune.Point(6, 2)
    assert {(q.block, q.push) for q in tune.launch_points(p)} == {(b, k) for b in (64, 128, 256) for k in (1, 2)}
    assert [q.chunk for q in tune.chunk_points(p)] == [1, 4, 16, 64]
    occ = tune.occupancy_points(p)
    assert [q.waves_per_cu for q in occ[:4]] == [4, 8, 12, 16] and [q.stack_extra for q in occ[4:]] == [1, 2]
    # The configuration timed for a point: stack depth for the depth-`preplace` tasks (the count
    # takes the split rows off it).
    c = tune.Point(6, 2, block=256, chunk=4, push=2, waves_per_cu=12, stack_extra=1).config(21)
    assert (c.stack_depth, c.split, c.block, c.chunk, c.push, c.waves_per_cu) == (15, 2, 256, 4, 2, 12)


def test_tuner_cache_file_format(tmp_path):
    """The cache is plain text the C library parses (nq_tune.cpp): the last line for (arch, N)
    wins, comments and malformed lines are skipped."""
    from n_queens_amd import tune
    p = tmp_path / "t.txt"
    p.write_text("# comment\n"
                 "gfx950:sramecc+:xnack- N=21 preplace=6 split=2 block=256 chunk=1 push=2 waves_per_cu=0 "
                 "stack_extra=0 ms=7345.1\n"
                 "gfx950:sramecc+:xnack- N=21 preplace=5 split=3 block=128 chunk=4 push=1 waves_per_cu=12 "
                 "stack_extra=1 ms=7300.0\n"
                 "gfx950:sramecc+:xnack- N=20 preplace=6 broken\n")
    c = tune.load_cache(str(p))
    assert c[("gfx950:sramecc+:xnack-", 21)]["preplace"] == 5 and c[("gfx950:sramecc+:xnack-", 21)]["ms"] == 7300.0
    assert ("gfx950:sramecc+:xnack-", 20) not in c or "split" not in c[("gfx950:sramecc+:xnack-", 20)]


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


def test_default_and_solve_configs_leave_block_automatic():
    """nq_config_default / nq_solve_config need no device: workgroup size 0 (automatic, chosen by
    occupancy at launch), the smallest legal stack, chunk 16, automatic push, and the split rule
    clamp(N - rows - R, 0, 8) with R = (N+4)/2 up to N=20, 12 at N=21 and 13 from N=22 on."""
    c = nq.config_default(21, 9)
    assert (c.block, c.stack_depth, c.chunk, c.push, c.split) == (0, 11, 16, nq.NQ_PUSH_AUTO, 0)
    for N, rows, split in ((21, 6, 3), (22, 6, 3), (20, 6, 2), (19, 6, 2), (18, 6, 1), (17, 5, 2), (16, 4, 2),
                           (13, 5, 0), (23, 6, 4), (21, 4, 5), (24, 6, 5), (27, 7, 7), (28, 7, 8), (28, 2, 8), (24, 10, 1)):
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


def test_tuner_cache_parsed_by_the_library_reader(tmp_path):
    """The C reader of the tuner's cache (n_queens_amd/csrc/nq_tune.cpp), built on the host with a
    small driver: the last entry for (arch, N) wins, timing fields are ignored, comments and
    malformed or out-of-range entries are skipped, and NQ_TUNE_CACHE="" switches the cache off."""
    drv = tmp_path / "drv.cpp"
    drv.write_text('#include <cstdio>\n#include <cstdlib>\n#include "nq_internal.h"\n'
                   "int main(int c, char **v) { nq::TunedEntry e; bool ok = nq::find_tuned(v[1], atoi(v[2]), &e);\n"
                   ' printf("%d %d %d %d %d %d %d %d %d\\n", ok, e.N, e.preplace, e.split, e.block, e.chunk, e.push,'
                   " e.waves_per_cu, e.stack_extra); }\n")
    exe = tmp_path / "drv"
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                           os.path.join(ROOT, "n_queens_amd", "csrc"), str(drv),
                           os.path.join(ROOT, "n_queens_amd", "csrc", "nq_tune.cpp"), "-o", str(exe)])
    cache = tmp_path / "t.txt"
    cache.write_text("# tuned on one device\n"
                     "gfx950:a N=17 preplace=7 split=2 block=64 chunk=4 push=1 waves_per_cu=8 stack_extra=1 ms=3.5\n"
                     "gfx950:a N=21 preplace=6 split=2 block=256 chunk=1 push=2 waves_per_cu=0 stack_extra=0\n"
                     "gfx950:a N=21 preplace=7 split=1 block=0 chunk=0 push=0 waves_per_cu=0 stack_extra=0 ms=7336.0\n"
                     "gfx950:a N=16 preplace=x split=2\n"
                     "gfx950:a N=15 preplace=5 split=9 block=0 chunk=0 push=0 waves_per_cu=0 stack_extra=0\n"
                     "gfx942:b N=17 preplace=5 split=1 block=0 chunk=0 push=0 waves_per_cu=0 stack_extra=0\n")

    def look(arch, N, path=str(cache)):
        out = subprocess.check_output([str(exe), arch, str(N)], env=dict(os.environ, NQ_TUNE_CACHE=path))
        return [int(x) for x in out.split()]

    assert look("gfx950:a", 17) == [1, 17, 7, 2, 64, 4, 1, 8, 1]
    assert look("gfx950:a", 21) == [1, 21, 7, 1, 0, 0, 0, 0, 0]   # the last line wins
    assert look("gfx950:a", 16)[0] == 0 and look("gfx950:a", 15)[0] == 0  # malformed, split > 4
    assert look("gfx942:b", 17)[:4] == [1, 17, 5, 1] and look("gfx950:a", 18)[0] == 0
    assert look("gfx950:a", 17, path="")[0] == 0


def _expand_plain(N, c, l, r, rows):
    """Every placement `rows` rows below (c, l, r), depth-first, lowest column first."""
    last = (1 << N) - 1
    out = []

    def rec(c, l, r, k):
        if k == 0:
            out.append((c, l & 0xFFFFFFFF, r))
            return
        v = last & ~(c | l | r)
        while v:
            p = v & -v
            v ^= p
            rec(c | p, ((l | p) << 1) & 0xFFFFFFFF, (r | p) >> 1, k - 1)

    rec(c, l, r, rows)
    return np.array(out, dtype=np.uint32).reshape(-1, 3)


def test_instance_expansion_equals_the_oracle(tmp_path):
    """The 16-byte task lists of the batching and serving paths (append_instance_tasks: the
    threaded two-pass expander of n_queens_amd/csrc/nq_prefix.cpp), built on the host with a small
    driver: a whole board is the oracle's mirror-symmetric task list (tests also the threaded
    path, 7 rows) with weight bit and tag; a prefixed instance is every placement below it in
    depth-first order (6 rows: threaded as well)."""
    drv = tmp_path / "drv.cpp"
    drv.write_text('#include <cstdio>\n#include <cstdlib>\n#include "nq_tasks.h"\n'
                   "int main(int argc, char **v) {\n"
                   " int N = atoi(v[1]), dt = atoi(v[2]), k = argc - 4; uint32_t c = 0, l = 0, r = 0;\n"
                   " for (int i = 0; i < k; ++i) { uint32_t p = 1u << atoi(v[4 + i]); c |= p; l = (l | p) << 1;"
                   " r = (r | p) >> 1; }\n"
                   " nq::TaskVec t; int d = nq::append_instance_tasks(N, k, c, l, r, dt, 5, t);\n"
                   " FILE *f = fopen(v[3], \"wb\"); fwrite(t.data(), 16, t.size(), f); fclose(f);\n"
                   ' printf("%d\\n", d); }\n')
    exe = tmp_path / "drv"
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I",
                           os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "n_queens_amd", "csrc"), str(drv),
                           "-L", os.path.join(ROOT, "n_queens_amd"), "-lnq_hip",
                           "-Wl,-rpath," + os.path.join(ROOT, "n_queens_amd"), "-o", str(exe)])

    def run(N, dt, cols=()):
        out = tmp_path / "t.bin"
        d = int(subprocess.check_output([str(exe), str(N), str(dt), str(out), *map(str, cols)], env=NO_GPU))
        return d, np.fromfile(out, dtype=np.uint32).reshape(-1, 4)

    for N, dt in ((5, 12), (8, 5), (12, 8), (17, 10)):
        rows = max(2, N - dt)
        d, t = run(N, dt)
        want = oracle.gen_tasks(N, rows)
        assert d == rows and np.array_equal(t[:, :3], want), (N, dt)
        assert np.all(t[:, 3] == (5 << 6) | (1 << 5) | N)
    for N, dt, cols in ((10, 6, [0, 2]), (16, 8, [1, 3]), (13, 20, [4])):
        enc = oracle.encode_prefix(N, cols)
        s = max(0, N - len(cols) - dt)
        d, t = run(N, dt, cols)
        assert d == len(cols) + s
        assert np.array_equal(t[:, :3], _expand_plain(N, *enc, s)), (N, cols)
        assert np.all(t[:, 3] == (5 << 6) | N)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and exits before touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_cli_checkpoint_refusals_without_gpu(tmp_path):
    """`--checkpoint` counts one range per call of nq_solve_resumable: a checkpoint of another
    problem is refused before any device work, and with no device the call fails after writing
    the initial checkpoint; either way the CLI prints the library's message and result -1, and no
    progress line."""
    foreign = tmp_path / "foreign"
    foreign.write_text("nq-checkpoint 1\nN 14 rows 4 ntasks 4816 next 0 sum 0\n")
    p = run_cli(13, 5, "--checkpoint", str(foreign))
    assert "checkpoint is for a different problem" in p.stderr, p.stderr
    assert re.search(r"cuda 13 queens result -1, calc time:", p.stdout) and "tasks" not in p.stdout
    fresh = tmp_path / "fresh"
    p = run_cli(13, 5, "--checkpoint", str(fresh), "--range-tasks", "1000")
    assert re.search(r"cuda 13 queens result -1, calc time:", p.stdout) and "done" not in p.stdout
    assert fresh.read_text() == "nq-checkpoint 1\nN 13 rows 5 ntasks 15550 next 0 sum 0\n"
