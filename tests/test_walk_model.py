"""CPU model of the device split in registers (split_kernel): the kernel's own TaskWalk and
LaneWalk (n_queens_amd/csrc/nq_walk.h) inside an emulation of its claim/refill control flow
(tests/walk_model.cpp).  For every case the model must claim each task exactly once, hand out
only roots that fit the configured stack, terminate, and reach the oracle's total."""
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = tmp_path_factory.mktemp("walk") / "walk_model"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "n_queens_amd", "csrc"),
                           os.path.join(ROOT, "tests", "walk_model.cpp"), "-o", str(exe)])
    return str(exe)


def run_model(model, tmp_path, N, tasks, split, chunk, waves, stack, lanes=False):
    f = tmp_path / "tasks.bin"
    np.ascontiguousarray(tasks, dtype=np.uint32).tofile(f)
    out = subprocess.check_output([model, str(N), str(split), str(chunk), str(waves), str(stack), str(f)]
                                  + (["lanes"] if lanes else []), timeout=120).split()
    total, flags, roots, claims_ok, frames_ok, done, cursor_claims, nodes = (int(x) for x in out)
    run_model.cursor_claims = cursor_claims
    run_model.nodes = nodes
    return total, flags, roots, claims_ok == 1, frames_ok == 1, done == 1


@pytest.mark.parametrize("N", range(5, 14))
def test_walk_model_totals(model, tmp_path, N):
    for rows, split in itertools.product(range(1, N - 1), range(1, 9)):
        tasks = oracle.gen_tasks(N, rows)
        want = oracle.count_tasks(N, tasks)
        stack = max(1, N - 1 - (rows + split))
        for chunk, waves in ((1, 1), (3, 7), (16, 3)):
            total, flags, roots, claims, frames, done = run_model(model, tmp_path, N, tasks, split, chunk, waves,
                                                                  stack)
            assert (claims, frames, done, flags) == (True, True, True, 0), (N, rows, split, chunk, waves)
            assert total == want, (N, rows, split, chunk, waves, total, want)


def test_walk_model_mixed_depths_and_edge_tasks(model, tmp_path):
    """Tasks of every depth in one list (full boards, last-row prefixes, dead prefixes, shallow
    ones), with the stack sized for the shallowest: the walk counts the deep ones directly."""
    N = 11
    rng = np.random.default_rng(5)
    parts = [oracle.gen_tasks(N, d) for d in range(1, N + 1)]
    tasks = np.concatenate([p[rng.choice(len(p), min(len(p), 40), replace=False)] for p in parts if len(p)])
    tasks = np.concatenate([tasks, np.array([[0b11, 0b1100, 0], [0, 0, 0]], dtype=np.uint32)])  # dead, empty
    rng.shuffle(tasks)
    want = oracle.count_tasks(N, tasks)
    for split, lanes in itertools.chain(((s, False) for s in range(1, 9)), ((s, True) for s in range(1, 5))):
        stack = N - 1 - split  # the empty board (depth 0) is the shallowest
        total, flags, _, claims, frames, done = run_model(model, tmp_path, N, tasks, split, 2, 5, stack, lanes)
        assert (claims, frames, done, flags, total) == (True, True, True, 0, want), (split, lanes)


@pytest.mark.parametrize("N", [9, 12])
def test_lane_walk_equals_register_walk(model, tmp_path, N):
    """The walk with its frames in lanes (split_kernel's walk beyond 4 rows) run at 1..4 rows,
    where the scalar-register walk is the reference: the same totals, roots and claims."""
    for rows, split in itertools.product(range(1, N - 2), range(1, 5)):
        tasks = oracle.gen_tasks(N, rows)
        stack = max(1, N - 1 - (rows + split))
        a = run_model(model, tmp_path, N, tasks, split, 3, 7, stack)
        b = run_model(model, tmp_path, N, tasks, split, 3, 7, stack, lanes=True)
        assert a == b and a[0] == oracle.count_tasks(N, tasks), (N, rows, split)


def test_walk_model_guided_claims_are_few(model, tmp_path):
    """Guided claims: a wave takes a share of the tasks left, so a long list costs few cursor
    atomics (each is device-wide traffic), and single tasks near the end keep the tail short."""
    N, rows = 13, 5
    tasks = oracle.gen_tasks(N, rows)  # 15,550 tasks
    total, flags, _, claims, frames, done = run_model(model, tmp_path, N, tasks, 2, 1, 7, N - 1 - rows - 2)
    assert (claims, frames, done, flags, total) == (True, True, True, 0, oracle.count_tasks(N, tasks))
    assert run_model.cursor_claims < len(tasks) // 20, run_model.cursor_claims


def test_walk_model_flags_a_stack_too_small(model, tmp_path):
    N, rows, split = 12, 3, 2
    tasks = oracle.gen_tasks(N, rows)
    total, flags, roots, claims, frames, done = run_model(model, tmp_path, N, tasks, split, 4, 3,
                                                          N - 1 - (rows + split) - 1)
    assert flags == 1 and roots == 0 and claims and done


@pytest.mark.parametrize("lanes", [False, True])
def test_walk_model_per_task_counts(model, tmp_path, lanes):
    """split_kernel's per-task counts (PT): lanes add their roots' counts per task when they move
    to another task and at exit, the walk's last-row counts go to their task when it moves on, and
    edge tasks' when claimed; every task's count equals the oracle's, for plain lists and a list of
    mixed depths, at every split depth and several claim sizes and wave counts."""
    cases = []
    for N, rows in ((9, 2), (11, 3), (12, 4), (13, 5)):
        cases.append((N, oracle.gen_tasks(N, rows), rows))
    N = 11
    rng = np.random.default_rng(7)
    parts = [oracle.gen_tasks(N, d) for d in range(1, N + 1)]
    mixed = np.concatenate([p[rng.choice(len(p), min(len(p), 30), replace=False)] for p in parts if len(p)])
    mixed = np.concatenate([mixed, np.array([[0b11, 0b1100, 0], [0, 0, 0]], dtype=np.uint32)])
    rng.shuffle(mixed)
    cases.append((N, mixed, 0))
    out = tmp_path / "per_task.bin"
    for N, tasks, d0 in cases:
        _, want = oracle.count_tasks(N, tasks, per_task=True)
        for split in range(1, 9):
            if d0 + split > N - 2:
                continue
            for chunk, waves in ((1, 1), (3, 7)):
                f = tmp_path / "tasks.bin"
                np.ascontiguousarray(tasks, dtype=np.uint32).tofile(f)
                stack = max(1, N - 1 - (d0 + split))
                res = subprocess.check_output([model, str(N), str(split), str(chunk), str(waves), str(stack), str(f),
                                               "lanes" if lanes else "regs", str(out)], timeout=120).split()
                assert int(res[1]) == 0 and int(res[0]) == int(want.sum()), (N, split, chunk, waves)
                got = np.fromfile(out, dtype=np.uint64)
                assert np.array_equal(got, want), (N, d0, split, chunk, waves)


@pytest.mark.parametrize("lanes", [False, True])
def test_walk_model_node_counts(model, tmp_path, lanes):
    """split_kernel's node count (MODE 2): the placements the walk makes plus, for every root, the
    hot loop's expansions less its completed boards, equals the oracle's kernel-sense node count
    (placements at depths rows+1..N-1) at every split depth, for plain and mixed-depth lists."""
    N = 11
    rng = np.random.default_rng(11)
    parts = [oracle.gen_tasks(N, d) for d in range(1, N + 1)]
    mixed = np.concatenate([p[rng.choice(len(p), min(len(p), 30), replace=False)] for p in parts if len(p)])
    cases = [(9, oracle.gen_tasks(9, 2), 2), (12, oracle.gen_tasks(12, 3), 3), (13, oracle.gen_tasks(13, 5), 5),
             (N, mixed, 1)]
    for n, tasks, d0 in cases:
        want = oracle.count_tasks_nodes(n, tasks)
        for split in range(1, 9):
            if d0 + split > n - 2:
                continue
            stack = max(1, n - 1 - (d0 + split))
            f = tmp_path / "tasks.bin"
            np.ascontiguousarray(tasks, dtype=np.uint32).tofile(f)
            out = subprocess.check_output([model, str(n), str(split), "2", "5", str(stack), str(f),
                                           "lanes" if lanes else "regs"], timeout=120).split()
            assert int(out[1]) == 0 and int(out[7]) == want, (n, d0, split, int(out[7]), want)
