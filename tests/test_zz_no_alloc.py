"""nq_count_static, the entry point that allocates no device memory, against nq_count through the
C ABI: two calls in a row on the same device (the second with a smaller task list) must give
nq_count's sum, per-task counts and node counts bit for bit; the split path too; a list beyond
the static scratch is refused before any device work."""
import numpy as np
import pytest

import n_queens_amd as nq
import oracle

pytestmark = pytest.mark.gpu


def _check_against_count(N, tasks, **kw):
    a = nq.count_static(N, tasks, **kw)
    b = nq.count(N, tasks, **kw)
    assert a.sum == b.sum, (N, kw)
    if kw.get("per_task"):
        assert np.array_equal(a.per_task, b.per_task), (N, kw)
    if kw.get("nodes"):
        assert a.nodes == b.nodes, (N, kw)
    return a


def test_static_twice_second_smaller(published):
    big = nq.generate(16, 5)      # 58,304 tasks
    small = nq.generate(13, 4)    # smaller list, other N: the scratch keeps the first call's data
    a = _check_against_count(16, big, per_task=True, nodes=True)
    assert 2 * a.sum == published["solutions"][16]
    b = _check_against_count(13, small, per_task=True, nodes=True)
    assert np.array_equal(b.per_task, oracle.count_tasks(13, small, per_task=True)[1])
    assert 2 * b.sum == published["solutions"][13]


def test_static_totals_with_split_twice(published):
    for N, rows, split in ((18, 5, 2), (15, 4, 3)):
        t = nq.generate(N, rows)
        r = _check_against_count(N, t, cfg=nq.Config(split=split))
        assert 2 * r.sum == published["solutions"][N], (N, rows, split)
    r = _check_against_count(14, nq.generate(14, 3))  # automatic split
    assert 2 * r.sum == published["solutions"][14]


def test_static_edge_lists():
    assert nq.count_static(8, np.zeros((0, 3), np.uint32)).sum == 0
    assert nq.count_static(10, np.zeros((1, 3), np.uint32), per_task=True).per_task[0] == 724


def test_static_refuses_lists_beyond_its_scratch():
    t = np.zeros((nq.NQ_STATIC_MAX_TASKS + 1, 3), np.uint32)
    with pytest.raises(nq.NQError) as e:
        nq.count_static(12, t)
    assert e.value.status == nq.NQ_ERR_ARG and "scratch" in str(e.value)
