"""The batch plan's order (largest estimated task first) rests on est_nodes_per_task
(nq_batch.cpp), Knuth's random-probe estimate of a prefix's subtree.  Checked here on the CPU
against exact per-task node counts of the 512-instance mix: the oracle's node count of every
instance (tests/golden/mix_nodes.json) divided by its task count after the split to 12 free
rows, for the instances whose task count is cheap to enumerate here."""
import ctypes
import json
import os

import pytest

import n_queens_amd as nq
import oracle
from n_queens_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEPTH_TARGET = 12  # free rows a batch task keeps (nq_batch_plan_create's default)


@pytest.fixture(scope="module")
def est():
    f = nq.lib()["_ZN2nq18est_nodes_per_taskEiijjji"]  # nq::est_nodes_per_task(int, int, u32, u32, u32, int)
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    return f


def _tasks_below(N, c, l, r, rows):
    """Number of placements `rows` rows below a prefix (the batch's split)."""
    last = (1 << N) - 1
    front = [(c, l, r)]
    for _ in range(rows):
        nxt = []
        for c0, l0, r0 in front:
            v = last & ~(c0 | l0 | r0)
            while v:
                p = v & -v
                v ^= p
                nxt.append((c0 | p, ((l0 | p) << 1) & 0xFFFFFFFF, (r0 | p) >> 1))
        front = nxt
    return len(front)


@pytest.fixture(scope="module")
def truth():
    """(N, k, prefix, split, exact nodes per task) for instances split by at most 3 rows."""
    with open(os.path.join(ROOT, "tests", "golden", "mix_nodes.json")) as f:
        nodes = json.load(f)["nodes"]
    out = []
    for (N, cols), n in zip(workloads.batch_mix(), nodes):
        k = len(cols)
        split = max(0, N - k - DEPTH_TARGET)
        if split > 3:
            continue
        c, l, r = oracle.encode_prefix(N, cols)
        out.append((N, k, (c, l, r), split, n / _tasks_below(N, c, l, r, split)))
    return out


def test_estimate_is_deterministic(est):
    c, l, r = oracle.encode_prefix(16, [1, 3])
    assert est(16, 2, c, l, r, 2) == est(16, 2, c, l, r, 2) > 0


def test_estimate_within_band_of_exact(est, truth):
    assert len(truth) >= 300
    ratios = sorted(est(N, k, *pre, s) / t for N, k, pre, s, t in truth)
    p5, p95 = ratios[len(ratios) // 20], ratios[-len(ratios) // 20 - 1]
    assert 0.6 < p5 and p95 < 1.6, (p5, p95)


def test_order_puts_large_tasks_first(est, truth):
    """Sorted by the estimate, no instance in the last half of the order has tasks of more
    than a quarter of the largest tasks' size."""
    rows = sorted(truth, key=lambda x: -est(x[0], x[1], *x[2], x[3]))
    biggest = max(t for *_, t in rows)
    assert max(t for *_, t in rows[len(rows) // 2:]) < biggest / 4


def test_estimate_of_whole_small_boards(est):
    """With no split, a whole board's estimate is close to its exact DFS node count (the
    oracle's), and a prefix with no candidate below it has none."""
    for N in (8, 10, 12):
        exact = oracle.count_nodes(N, 0, 0, 0)
        assert 0.5 < est(N, 0, 0, 0, 0, 0) / exact < 2.0, N
    c, l, r = oracle.encode_prefix(4, [0, 2])  # every square of the third row is attacked
    assert oracle.count_nodes(4, c, l, r) == 0
    assert est(4, 2, c, l, r, 0) == 0
