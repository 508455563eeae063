"""Batched instances: validation on the CPU; GPU counts against the plain per-instance path,
the oracle and the reference's published totals."""
import itertools
import os

import numpy as np
import pytest

import n_queens_amd as nq
import oracle
from n_queens_amd import workloads

NO_GPU = os.environ.get("HIP_VISIBLE_DEVICES") == "-1"


def test_mix_generator_is_seeded_and_valid():
    a = workloads.batch_mix()
    b = workloads.batch_mix()
    assert a == b and len(a) == 512
    for N, cols in a:
        assert 14 <= N <= 20 and 2 <= len(cols) <= 6
        assert oracle.encode_prefix(N, cols) is not None


@pytest.mark.parametrize("bad", [(8, [0, 1]), (8, [0, 8]), (29, [0]), (5, [0, 2, 4, 1, 3, 0])])
def test_batch_rejects_invalid_instances(bad):
    # Validation happens before any device work, so this also runs without a GPU.
    with pytest.raises(nq.NQError) as e:
        nq.batch_count([(10, [0, 2]), bad])
    assert e.value.status == nq.NQ_ERR_ARG
    assert "instance 1" in str(e.value)


@pytest.mark.gpu
def test_batch_partition_sums_to_published(published):
    """All valid first-two-row prefixes of N=14 partition the board: counts sum to A000170."""
    N = 14
    insts = [(N, [a, b]) for a, b in itertools.product(range(N), repeat=2) if abs(a - b) > 1]
    counts, st = nq.batch_count(insts)
    assert int(counts.sum()) == published["solutions"][N]
    for (n, cols), got in zip(insts, counts):
        assert got == oracle.count_subtree(n, *oracle.encode_prefix(n, cols)), cols


@pytest.mark.gpu
def test_batch_mixed_sizes_and_edges(published):
    insts = [(N, []) for N in range(4, 13)]                       # whole boards, no fold
    insts += [(8, [0, 4, 7, 5, 2, 6, 1, 3]), (8, [0, 4, 7, 5, 2, 6, 1]), (8, [0, 2, 4]), (1, []),
              (27, [0, 2, 4, 1, 3, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 5, 7, 9, 11]), (20, [1, 3])]
    counts, _ = nq.batch_count(insts)
    for N in range(4, 13):
        assert counts[N - 4] == published["solutions"][N]
    for (N, cols), got in zip(insts[9:], counts[9:]):
        enc = oracle.encode_prefix(N, cols)
        assert enc is not None
        if N <= 20 and len(cols) >= 2 and N - len(cols) <= 18:
            assert got == oracle.count_subtree(N, *enc), (N, cols)


@pytest.mark.gpu
def test_batch_mix_vs_per_instance_launches():
    mix = workloads.batch_mix()
    counts, st = nq.batch_count(mix)
    per = np.array([nq.instance_count(N, cols) for N, cols in mix], dtype=np.uint64)
    assert np.array_equal(counts, per)
    # the oracle pins the instances it can count quickly
    for (N, cols), got in zip(mix, counts):
        if N <= 16:
            assert got == oracle.count_subtree(N, *oracle.encode_prefix(N, cols)), (N, cols)


@pytest.mark.gpu
def test_batch_plan_replay_is_stable():
    """Replays of the captured graph, with eager device work between them, each give the
    eager (un-captured) counts: the graph zeroes its counters with a kernel, not a memset."""
    mix = workloads.batch_mix(64, seed=3)
    eager, _ = nq.batch_count(mix)
    plan = nq.BatchPlan(mix)
    t13 = nq.generate(13, 5)
    for i in range(4):
        c, st = plan.run()
        assert np.array_equal(c, eager), i
        assert st["total"] == int(eager.sum())
        assert nq.count(13, t13).sum == 36856  # unrelated eager work between replays
    plan.launch()
    c, st = plan.wait()
    plan.close()
    assert np.array_equal(c, eager)


def _fixture(name):
    import json
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "instance_counts.json")) as f:
        fx = json.load(f)[name]
    return fx


@pytest.mark.gpu
def test_batch_mix_every_instance_vs_oracle_fixture():
    """The 512-instance mix (N up to 20): the batched launch, its graph replay and the per-instance
    launches all equal the oracle's per-instance counts (tests/golden/instance_counts.json)."""
    mix = workloads.batch_mix()
    fx = _fixture("batch_mix")
    assert fx["digest"] == workloads.instances_digest(mix)
    counts, _ = nq.batch_count(mix)
    assert [int(x) for x in counts] == fx["counts"]
    plan = nq.BatchPlan(mix)
    for _ in range(2):
        again, _ = plan.run()
        assert [int(x) for x in again] == fx["counts"]
    plan.close()
    per = [nq.instance_count(N, cols) for N, cols in mix]
    assert per == fx["counts"]
