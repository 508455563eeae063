"""GPU parity: the HIP DFS (through the C ABI) against the oracle and the reference's numbers.

All counts are integers and must match bit for bit (tolerance 0).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import n_queens_amd as nq
import oracle

pytestmark = pytest.mark.gpu


def test_device_visible():
    assert nq.device_count() >= 1
    info = nq.device_info(0)
    print(info)
    assert "gfx950" in info["name"]


def test_n13_rows5_per_task_vs_reference_vectors(ref_vectors):
    """P1: N=13, 5 rows pre-placed: every per-task count equals the reference CPU's."""
    t = ref_vectors["N13_r5_tasks"]
    r = nq.count(13, t, per_task=True)
    assert np.array_equal(r.per_task, ref_vectors["N13_r5_counts"])
    assert 2 * r.sum == 73712


def test_all_reference_vector_cases(ref_vectors):
    for N, rows, total in ref_vectors["totals"]:
        N, rows, total = int(N), int(rows), int(total)
        r = nq.count(N, ref_vectors[f"N{N}_r{rows}_tasks"], per_task=True)
        assert np.array_equal(r.per_task, ref_vectors[f"N{N}_r{rows}_counts"]), (N, rows)
        assert 2 * r.sum == total


@pytest.mark.parametrize("N", range(4, 17))
def test_totals_every_legal_rows(N, published):
    for rows in range(2, N - 1):
        total, _ = nq.solve(N, rows)
        assert total == published["solutions"][N], (N, rows)


@pytest.mark.parametrize("N", [16, 17, 18])
def test_p2_rows5_vs_published_and_oracle_sample(N, published):
    total, st = nq.solve(N, 5)
    assert total == published["solutions"][N]
    # per-task parity on a seeded sample of the tasks against the oracle
    t = nq.generate(N, 5)
    rng = np.random.default_rng(N)
    idx = np.sort(rng.choice(len(t), size=min(300, len(t)), replace=False))
    r = nq.count(N, t[idx], per_task=True)
    _, expect = oracle.count_tasks(N, t[idx], per_task=True)
    assert np.array_equal(r.per_task, expect)


@pytest.mark.parametrize("N,rows", [(12, 3), (13, 5), (14, 4), (16, 4), (17, 5)])
def test_node_count_vs_oracle(N, rows):
    t = nq.generate(N, rows)
    r = nq.count(N, t, nodes=True)
    assert r.nodes == oracle.count_tasks_nodes(N, t)
    assert r.sum == oracle.count_tasks(N, t)


@pytest.mark.parametrize("cfg", [
    nq.Config(block=64, chunk=1), nq.Config(block=128, chunk=8), nq.Config(block=256, chunk=64),
    nq.Config(block=64, stack_depth=24), nq.Config(block=192, chunk=3, grid=7), nq.Config(waves_per_cu=4),
    nq.Config(push=nq.NQ_PUSH_MASKED), nq.Config(push=nq.NQ_PUSH_ALWAYS, chunk=5),
    nq.Config(push=nq.NQ_PUSH_ALWAYS, block=256, waves_per_cu=8),
])
def test_launch_configs_same_counts(cfg, ref_vectors):
    t = ref_vectors["N13_r5_tasks"]
    r = nq.count(13, t, cfg=cfg, per_task=True)
    assert np.array_equal(r.per_task, ref_vectors["N13_r5_counts"]), cfg


def test_edge_tasks():
    """Empty list, dead prefixes, prefixes at depth N-1 and N, duplicates, mixed depths."""
    assert nq.count(8, np.zeros((0, 3), np.uint32)).sum == 0
    N = 8
    rows = []
    # dead-ish short prefixes, a full board (depth N), a depth N-1 prefix, mixed depths
    for cols in ([0, 2], [0, 3], [0, 4, 7, 5, 2, 6, 1, 3], [0, 4, 7, 5, 2, 6, 1], [0, 2, 4],
                 [1, 3, 5, 7], [1, 3, 0, 6], [2, 4, 1, 7, 0]):
        enc = oracle.encode_prefix(N, cols)
        if enc is not None:
            rows.append(enc)
    t = np.array(rows + rows[:2], dtype=np.uint32)
    r = nq.count(N, t, per_task=True)
    _, expect = oracle.count_tasks(N, t, per_task=True)
    assert np.array_equal(r.per_task, expect)
    # the empty board as a single task (depth 0)
    r = nq.count(10, np.zeros((1, 3), np.uint32), per_task=True)
    assert r.per_task[0] == 724


def test_depth_too_small_is_an_error():
    t = nq.generate(14, 3)
    with pytest.raises(nq.NQError) as e:
        nq.count(14, t, cfg=nq.Config(stack_depth=5))
    assert e.value.status == nq.NQ_ERR_DEPTH


def test_repeat_calls_are_independent():
    """The reference's cursor is never reset (n_queens_cuda.cu:8); here every call restarts."""
    a, _ = nq.solve(12, 3)
    b, _ = nq.solve(12, 3)
    assert a == b == 14200


def test_n20_rows6(published):
    total, st = nq.solve(20, 6)
    print(st)
    assert st["ntasks"] == 2967560
    assert total == published["solutions"][20]


def test_cli_end_to_end():
    p = subprocess.run([nq.CLI_PATH, "13", "5"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0
    assert "Use " in p.stdout and "to generate 15550 subproblems!" in p.stdout
    assert "cuda 13 queens result 73712, calc time:" in p.stdout


@pytest.mark.parametrize("N,rows", [(10, 2), (13, 3), (16, 4), (17, 5)])
def test_device_split_keeps_totals(N, rows, published):
    """Tasks split on the device by 1..3 rows (walked in registers by the DFS kernel) give the
    same totals."""
    t = nq.generate(N, rows)
    base = nq.count(N, t).sum
    assert 2 * base == published["solutions"][N]
    for split in (1, 2, 3):
        r = nq.count(N, t, cfg=nq.Config(split=split))
        assert r.sum == base, (N, rows, split)
        total, _ = nq.solve(N, rows, nq.Config(split=split))
        assert total == published["solutions"][N]


def test_device_split_mixed_depth_list():
    """A split of a list holding full boards, depth N-1 and dead prefixes keeps every count."""
    N = 8
    rows = []
    for cols in ([0, 4, 7, 5, 2, 6, 1, 3], [0, 4, 7, 5, 2, 6, 1], [0, 2, 4], [1, 3, 5, 7], [2, 4, 1, 7, 0], [], [3]):
        rows.append(oracle.encode_prefix(N, cols))
    t = np.array(rows, dtype=np.uint32)
    expect = oracle.count_tasks(N, t)
    for split in (0, 1, 2, 4):
        assert nq.count(N, t, cfg=nq.Config(split=split)).sum == expect, split


@pytest.mark.parametrize("chunk", [2, 8, 64, 1000])
@pytest.mark.parametrize("split", [0, 1, 2])
def test_chunked_claims_with_device_split(chunk, split, published):
    """Chunked cursor claims stay exact, per lane (split 0: lanes retry the wave's LDS range
    when claimed tasks are dead) and per wave (split > 0: whole tasks per wave)."""
    for N, rows, block in ((17, 5, 64), (15, 3, 256), (12, 2, 128)):
        t = nq.generate(N, rows)
        r = nq.count(N, t, cfg=nq.Config(block=block, chunk=chunk, split=split))
        assert 2 * r.sum == published["solutions"][N], (N, rows, block, chunk, split)


@pytest.mark.parametrize("chunk", [3, 64])
def test_chunked_claims_per_task(chunk, ref_vectors):
    """Per-task counts (no split) with chunked claims: every task is counted exactly once."""
    N, rows = 13, 5
    t = nq.generate(N, rows)
    r = nq.count(N, t, per_task=True, cfg=nq.Config(block=64, chunk=chunk))
    assert np.array_equal(r.per_task, oracle.count_tasks(N, t, per_task=True)[1])


@pytest.mark.parametrize("chunk", [8, 64])
def test_n20_split2_chunked(chunk, published):
    total, _ = nq.solve(20, 6, nq.Config(block=64, chunk=chunk, split=2))
    assert total == published["solutions"][20]


@pytest.mark.parametrize("chunk", [4, 64])
def test_batch_chunked_claims(chunk):
    from n_queens_amd import workloads
    mix = workloads.batch_mix(96, seed=5)
    a, _ = nq.batch_count(mix, nq.Config(chunk=1))
    b, _ = nq.batch_count(mix, nq.Config(chunk=chunk))
    assert np.array_equal(a, b)
    for (N, cols), got in zip(mix, b):
        if N <= 15:
            assert got == oracle.count_subtree(N, *oracle.encode_prefix(N, cols)), (N, cols)


@pytest.mark.gpu
def test_automatic_block_and_resolved_launch(published):
    """Block 0 resolves to 256 when that keeps the resident waves of 128, else 128; explicit
    blocks are kept; counts do not depend on it."""
    big = 1 << 40
    cus = nq.device_info(0)["cus"]
    # N=21 after a 3-row split, 11 frames: the masked push keeps 14 waves/CU with 128 lanes (the
    # spare level of the unconditional push, or 256 lanes, would leave 12).
    r21 = nq.resolve_config(21, 9, big)
    assert (r21.block, r21.push, r21.stack_depth) == (128, nq.NQ_PUSH_MASKED, 11) and r21.grid * 2 == 14 * cus
    # 12 frames: 12 waves/CU with either push form or block, so the unconditional push and 256 lanes.
    r12 = nq.resolve_config(21, 8, big)
    assert (r12.block, r12.push) == (256, nq.NQ_PUSH_ALWAYS) and r12.grid * 4 == 12 * cus
    assert nq.resolve_config(21, 9, big, nq.Config(block=64)).block == 64
    assert nq.resolve_config(16, 7, 10).grid == 1  # grid capped by the lanes the tasks need
    t = nq.generate(16, 5)
    sums = {nq.count(16, t, cfg=nq.Config(block=b)).sum for b in (0, 128, 256)}
    assert len(sums) == 1 and 2 * sums.pop() == published["solutions"][16]


@pytest.mark.parametrize("push", [nq.NQ_PUSH_MASKED, nq.NQ_PUSH_ALWAYS])
def test_push_forms_exact(push, published):
    """Both DFS push forms (exec-masked and unconditional with a spare level) give the same
    totals, per-task counts and node counts, also where the stack is used to its last level."""
    for N, rows in ((14, 3), (16, 4), (18, 5)):
        total, _ = nq.solve(N, rows, nq.Config(push=push, split=-1))
        assert total == published["solutions"][N], (N, rows, push)
    N, rows = 13, 2
    t = nq.generate(N, rows)
    r = nq.count(N, t, cfg=nq.Config(push=push), per_task=True, nodes=True)
    assert np.array_equal(r.per_task, oracle.count_tasks(N, t, per_task=True)[1])
    assert r.nodes == oracle.count_tasks_nodes(N, t)
    # Tasks at depth 0 (the whole board): every lane may fill all N-1 levels.
    t0 = np.zeros((1, 3), np.uint32)
    for N in (13, 14, 15):
        assert nq.count(N, t0, cfg=nq.Config(push=push)).sum == published["solutions"][N]


def test_n28_deep_prefixes():
    """N = 28 (the reference's largest board): prefixes 16 rows deep, counted by nq_count and by
    the batched kernel, against the oracle."""
    N = 28
    rng = np.random.default_rng(28)
    prefixes = []
    while len(prefixes) < 64:
        cols = _random_prefix(rng, N, 16)
        if cols is not None:
            prefixes.append(cols)
    t = np.array([oracle.encode_prefix(N, cols) for cols in prefixes], dtype=np.uint32)
    expect = np.array([oracle.count_subtree(N, *x) for x in t], dtype=np.uint64)
    assert int(expect.sum()) > 0
    r = nq.count(N, t, per_task=True)
    assert np.array_equal(r.per_task, expect)
    counts, _ = nq.batch_count([(N, cols) for cols in prefixes])
    assert np.array_equal(counts, expect)


def _random_prefix(rng, N, k):
    """A random non-attacking placement of k queens in the first k rows (None if stuck)."""
    cols = []
    for row in range(k):
        ok = [c for c in range(N) if all(c != q and abs(c - q) != row - i for i, q in enumerate(cols))]
        if not ok:
            return None
        cols.append(int(rng.choice(ok)))
    return cols


@pytest.mark.parametrize("N,rows", [(8, 3), (9, 4), (12, 2), (13, 5), (14, 7), (15, 1), (16, 6), (11, 9),
                                    (10, 10), (17, 3)])
def test_device_generation_matches_host_generator(N, rows):
    """nq_solve generates the depth-`rows` task list on the device from the depth-2 seeds: the
    same task count as the reference's host generator, and the same 2 x sum."""
    tasks = nq.generate(N, rows)
    total, st = nq.solve(N, rows)
    assert st["ntasks"] == len(tasks), (N, rows)
    assert total == 2 * oracle.count_tasks(N, tasks), (N, rows)
    total0, st0 = nq.solve(N, rows, nq.Config(split=0))
    assert total0 == total and st0["ntasks"] == len(tasks)


def test_resumable_interrupt_and_resume(tmp_path, published):
    """Counting in ranges, stopping after 4 of them and resuming from the checkpoint file gives
    the uninterrupted total; a finished checkpoint returns the total without device work."""
    p = str(tmp_path / "ck")
    a = nq.solve_resumable(13, 5, p, range_tasks=1000, max_ranges=4)
    assert not a["done"] and a["next"] == 4000 and a["ranges"] == 4 and a["total"] == 0
    tasks = nq.generate(13, 5)
    assert a["sum"] == oracle.count_tasks(13, tasks[:4000])
    b = nq.solve_resumable(13, 5, p, range_tasks=3000)
    assert b["done"] and b["next"] == b["ntasks"] == len(tasks) and b["ranges"] == 4
    assert b["total"] == published["solutions"][13]
    c = nq.solve_resumable(13, 5, p)
    assert c["done"] and c["ranges"] == 0 and c["total"] == b["total"]
    d = nq.solve_resumable(16, 4, str(tmp_path / "ck16"))
    assert d["done"] and d["ranges"] == 100 and d["total"] == published["solutions"][16]


def test_cli_checkpoint_resume(tmp_path):
    """`n_queens N rows --checkpoint PATH`: one progress line per counted range, in task order;
    a finished checkpoint is reported by the next run."""
    p = str(tmp_path / "ck")
    r = subprocess.run([nq.CLI_PATH, "14", "4", "--checkpoint", p, "--range-tasks", "300"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "cuda 14 queens result 365596, calc time:" in r.stdout, r.stdout + r.stderr
    n = len(nq.generate(14, 4))
    ranges = (n + 299) // 300
    done = [int(m) for m in re.findall(r"checkpoint \S+: tasks (\d+) of %d done" % n, r.stdout)]
    assert done == [min(300 * (i + 1), n) for i in range(ranges)], r.stdout
    assert f"{ranges} ranges counted in this run, tasks {n} of {n} done" in r.stdout
    r2 = subprocess.run([nq.CLI_PATH, "14", "4", "--checkpoint", p], capture_output=True, text=True, timeout=120)
    assert "0 ranges counted in this run" in r2.stdout and "cuda 14 queens result 365596" in r2.stdout


def test_node_count_n18_vs_oracle_threads():
    """N=18 rows 6 (2.9e10 nodes): the GPU node kernel against the oracle's threaded counter."""
    N, rows = 18, 6
    t = nq.generate(N, rows)
    r = nq.count(N, t, nodes=True)
    assert r.nodes == oracle.count_tasks_nodes(N, t, threads=16)
    assert 2 * r.sum == 666090624


def test_n21_node_count_fixture(published):
    """The N=21 rows 6 node count that bench.py's nodes/s is computed from."""
    N, rows = 21, 6
    t = nq.generate(N, rows)
    r = nq.count(N, t, nodes=True)
    assert r.nodes == published["dfs_nodes"]["21,6"]
    assert 2 * r.sum == published["solutions"][N]


@pytest.mark.gpu
@pytest.mark.parametrize("N,rows", [(16, 4), (18, 5), (20, 6)])
def test_solve_with_its_own_reported_config(N, rows, published, monkeypatch):
    """nq_solve given back what nq_solve_config reports (stack depth for the depth-`rows` tasks,
    the automatic split) counts exactly, as with no configuration (built-in rule: the tuner's
    cache is switched off)."""
    monkeypatch.setenv("NQ_TUNE_CACHE", "")
    cfg = nq.solve_config(N, rows)
    assert cfg.split > 0 and cfg.stack_depth == N - 1 - rows
    assert nq.solve(N, rows, cfg)[0] == published["solutions"][N]
    assert nq.solve(N, rows)[0] == published["solutions"][N]
