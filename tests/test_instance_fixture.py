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
