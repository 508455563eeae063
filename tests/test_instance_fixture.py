"""The per-instance fixture of the benchmark workloads (tests/golden/instance_counts.json, made by
tests/golden/make_instance_counts.py with the oracle): it belongs to the instances the seeded
generators produce, its whole boards are A000170, and its smaller instances agree with the
oracle and with an independent brute force here."""
import itertools
import json
import os

import pytest

import oracle
from n_queens_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "instance_counts.json")


@pytest.fixture(scope="module")
def fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def workload(name):
    if name == "batch_mix":
        return workloads.batch_mix()
    return [(N, cols) for _, N, cols in workloads.poisson_requests(400, lam=1.0)]


@pytest.mark.parametrize("name", ["batch_mix", "poisson"])
def test_fixture_belongs_to_the_seeded_workload(fixture, name, published):
    insts = workload(name)
    fx = fixture[name]
    assert fx["digest"] == workloads.instances_digest(insts)
    assert len(fx["counts"]) == len(insts) and sum(fx["counts"]) == fx["total"]
    for (N, cols), c in zip(insts, fx["counts"]):
        if not cols:
            assert c == published["solutions"][N], N


@pytest.mark.parametrize("name", ["batch_mix", "poisson"])
def test_fixture_small_instances_vs_oracle(fixture, name):
    """Every instance with at most 13 free rows is recounted here by the oracle (seconds)."""
    insts = workload(name)
    checked = 0
    for (N, cols), c in zip(insts, fixture[name]["counts"]):
        if cols and N - len(cols) <= 13:
            assert c == oracle.count_subtree(N, *oracle.encode_prefix(N, cols)), (N, cols)
            checked += 1
    assert checked >= 50


def test_fixture_tiny_instances_vs_brute_force(fixture):
    """Instances with at most 8 free rows, by permutations of the free columns (no bitmasks)."""
    checked = 0
    for name in ("batch_mix", "poisson"):
        for (N, cols), c in zip(workload(name), fixture[name]["counts"]):
            if cols and N - len(cols) <= 8:
                free = [x for x in range(N) if x not in cols]
                n = 0
                for rest in itertools.permutations(free):
                    q = [*cols, *rest]
                    if len({i + x for i, x in enumerate(q)}) == N and len({i - x for i, x in enumerate(q)}) == N:
                        n += 1
                assert c == n, (N, cols)
                checked += 1
    assert checked >= 5


MIX_NODES = os.path.join(ROOT, "tests", "golden", "mix_nodes.json")


@pytest.fixture(scope="module")
def mix_nodes():
    with open(MIX_NODES) as f:
        return json.load(f)


def test_mix_nodes_belong_to_the_seeded_mix(mix_nodes):
    """The node fixture (tests/golden/make_mix_nodes.py) is for the 512 seeded instances."""
    mix = workloads.batch_mix()
    assert mix_nodes["digest"] == workloads.instances_digest(mix)
    assert len(mix_nodes["nodes"]) == len(mix) == 512
    assert sum(mix_nodes["nodes"]) == mix_nodes["total"]


def _nodes_by_sets(N, cols):
    """DFS nodes below a prefix without bitmasks: queens placed row by row on columns and the
    two diagonal sets, counted at depths k+1..N-1 (a full board is not a node)."""
    used = set(cols)
    d1 = {i + x for i, x in enumerate(cols)}
    d2 = {i - x for i, x in enumerate(cols)}

    def walk(row):
        if row == N:
            return 0
        n = 0
        for x in range(N):
            if x in used or row + x in d1 or row - x in d2:
                continue
            if row < N - 1:
                n += 1
            used.add(x); d1.add(row + x); d2.add(row - x)
            n += walk(row + 1)
            used.discard(x); d1.discard(row + x); d2.discard(row - x)
        return n

    return walk(len(cols))


def test_mix_nodes_small_instances_by_sets(mix_nodes):
    """Instances with at most 9 free rows, counted with sets here (no bitmask code shared)."""
    checked = 0
    for (N, cols), n in zip(workloads.batch_mix(), mix_nodes["nodes"]):
        if N - len(cols) <= 9:
            assert n == _nodes_by_sets(N, cols), (N, cols)
            checked += 1
    assert checked >= 5


def test_mix_nodes_medium_instances_vs_oracle(mix_nodes):
    """Instances with at most 12 free rows, recounted by the oracle's node counter."""
    checked = 0
    for (N, cols), n in zip(workloads.batch_mix(), mix_nodes["nodes"]):
        if N - len(cols) <= 12:
            assert n == oracle.count_nodes(N, *oracle.encode_prefix(N, cols)), (N, cols)
            checked += 1
    assert checked >= 20
