"""GPU parity: the HIP DFS (through the C ABI) against the oracle and the reference's numbers.

All counts are integers and must match bit for bit (tolerance 0).
"""
import os
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


@pytest.mark.parametrize("N,rows", [(12, 3), (13, 5), (14, 4)])
def test_node_count_vs_oracle(N, rows):
    t = nq.generate(N, rows)
    r = nq.count(N, t, nodes=True)
    assert r.nodes == oracle.count_tasks_nodes(N, t)
    assert r.sum == oracle.count_tasks(N, t)


@pytest.mark.parametrize("cfg", [
    nq.Config(block=64, chunk=1), nq.Config(block=128, chunk=8), nq.Config(block=256, chunk=64),
    nq.Config(block=64, stack_depth=24), nq.Config(block=192, chunk=3, grid=7), nq.Config(waves_per_cu=4),
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
