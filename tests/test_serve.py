"""Serving with continuous batching: admission errors, results alone vs in the mix, Poisson
arrivals, empty-board requests."""
import os
import subprocess
import sys
import threading
import time

import numpy as np
import pytest

import n_queens_amd as nq
import oracle
from n_queens_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_server_without_device_is_an_error():
    code = ("import n_queens_amd as nq\n"
            "try:\n    nq.Server()\nexcept nq.NQError as e:\n    print('status', e.status)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1", PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120).stdout
    assert f"status {nq.NQ_ERR_NO_DEVICE}" in out


@pytest.mark.gpu
def test_served_counts_equal_alone_counts():
    reqs = workloads.poisson_requests(120, lam=1000.0, seed=11)
    alone = [nq.instance_count(N, cols) for _, N, cols in reqs]
    with nq.Server() as srv:
        tickets = [srv.submit(N, cols) for _, N, cols in reqs]
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
        st = srv.stats()
    assert got == alone
    assert st["completed"] == len(reqs)
    for (_, N, cols), c in zip(reqs, got):
        if N <= 14:
            assert c == oracle.count_subtree(N, *oracle.encode_prefix(N, cols)), (N, cols)


@pytest.mark.gpu
def test_requests_join_the_running_kernel():
    """Continuous batching into the persistent kernel: small requests submitted while a large one
    is on the device are published into the same ring, counted by the running kernel and finish
    long before the large one does; every count is exact."""
    big = (19, [])                          # ~2e11 DFS nodes: about 0.2 s on the device
    small = [(12, [0, 2]), (13, [1, 3]), (14, [0, 2, 4]), (12, [])]
    with nq.Server() as srv:
        tb = srv.submit(*big)
        t0 = time.perf_counter()            # the request is admitted (its latency counts from here)
        time.sleep(0.005)                   # the large request is running now
        ts = [(time.perf_counter(), srv.submit(N, cols)) for N, cols in small]
        got_small = [(t_sub, srv.wait(t, timeout_ms=60000)) for t_sub, t in ts]
        t_small_done = time.perf_counter()
        cb, sb = srv.wait(tb, timeout_ms=120000)
        t_big_done = t0 + sb["latency_ms"] / 1e3  # at most a few microseconds late
    assert cb == 4968057848
    for (N, cols), (_, (c, _)) in zip(small, got_small):
        want = 14200 if not cols else oracle.count_subtree(N, *oracle.encode_prefix(N, cols))
        assert c == want, (N, cols)
    # Every small request was submitted after the large one had started, and was done before it.
    assert all(t_sub > t0 for t_sub, _ in got_small)
    assert t_small_done < t_big_done, (t_small_done - t0, sb)
    assert max(s["latency_ms"] for _, (_, s) in got_small) < sb["latency_ms"] / 2


@pytest.mark.gpu
def test_poisson_arrivals_overlap_and_are_exact():
    reqs = workloads.poisson_requests(80, lam=400.0, seed=5)
    alone = [nq.instance_count(N, cols) for _, N, cols in reqs]
    tickets = [None] * len(reqs)
    with nq.Server() as srv:
        t0 = time.perf_counter()

        def feed():
            for i, (at, N, cols) in enumerate(reqs):
                delay = at - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(N, cols)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = [srv.wait(t, timeout_ms=60000) for t in tickets]
        st = srv.stats()
    assert [c for c, _ in res] == alone
    assert all(s["latency_ms"] >= s["queue_ms"] >= 0 for _, s in res)
    assert st["batches"] >= 2
    # Requests overlap on the device: some request was still in flight when a later one finished.
    spans = [(at, at + s["latency_ms"] / 1e3) for (at, _, _), (_, s) in zip(reqs, res)]
    assert any(a1 < a2 and e1 > e2 for a1, e1 in spans for a2, e2 in spans), "no two requests overlapped"


@pytest.mark.gpu
def test_empty_board_requests(published):
    with nq.Server() as srv:
        t = {N: srv.submit(N, []) for N in range(1, 17)}
        got = {N: srv.wait(x, timeout_ms=60000)[0] for N, x in t.items()}
    for N in range(4, 17):
        assert got[N] == published["solutions"][N]
    assert got[1] == 1 and got[2] == 0 and got[3] == 0


@pytest.mark.gpu
def test_admission_budget_and_invalid_requests():
    with nq.Server(max_pending_tasks=1000, slot_tasks=5000) as srv:   # a ring of 8192 tasks
        with pytest.raises(nq.NQError) as e:
            srv.submit(10, [0, 1])
        assert e.value.status == nq.NQ_ERR_ARG
        with pytest.raises(nq.NQError) as e:
            srv.submit(19, [])          # more tasks than the admission budget: never fits
        assert e.value.status == nq.NQ_ERR_NOMEM
        busy = 0
        tickets = []
        for _ in range(200):
            try:
                tickets.append(srv.submit(15, [0, 2]))
            except nq.NQError as err:
                assert err.status == nq.NQ_ERR_BUSY
                busy += 1
        counts = {srv.wait(t, timeout_ms=60000)[0] for t in tickets}
        assert counts == {oracle.count_subtree(15, *oracle.encode_prefix(15, [0, 2]))}


@pytest.mark.gpu
def test_request_larger_than_the_ring(published):
    """A whole N=17 board (about 1e5 tasks) through a ring of 8192 tasks: published slice by
    slice as ring space frees up, with the ring wrapping many times; the count is A000170's."""
    with nq.Server(slot_tasks=5000, max_pending_tasks=1 << 20) as srv:
        c, st = srv.wait(srv.submit(17, []), timeout_ms=60000)
        stats = srv.stats()
    assert st["ntasks"] > 8192 * 4
    assert c == published["solutions"][17]
    assert stats["batches"] >= st["ntasks"] // 8192


@pytest.mark.gpu
def test_served_poisson_set_vs_oracle_fixture():
    """The 400 Poisson requests bench.py serves (N up to 19, whole boards included), submitted at
    once: every served count equals the oracle's (tests/golden/instance_counts.json)."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "instance_counts.json")) as f:
        fx = json.load(f)["poisson"]
    reqs = [(N, cols) for _, N, cols in workloads.poisson_requests(400, lam=1.0)]
    assert fx["digest"] == workloads.instances_digest(reqs)
    with nq.Server() as srv:
        tickets = []
        for N, cols in reqs:
            for _ in range(60000):
                try:
                    tickets.append(srv.submit(N, cols))
                    break
                except nq.NQError as e:
                    assert e.status == nq.NQ_ERR_BUSY
                    time.sleep(0.001)
            else:
                raise AssertionError("request refused as busy for 60 s")
        got = [srv.wait(t, timeout_ms=120000)[0] for t in tickets]
    assert got == fx["counts"]


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [dict(chunk=1), dict(chunk=64, waves_per_cu=4), dict(depth_target=8),
                                 dict(depth_target=14, block=256), dict(slot_tasks=1000, slot_requests=3),
                                 dict(batch_work=1e6)])
def test_server_configurations_exact(cfg):
    """Served counts equal the alone counts across the server's configuration space: single-ticket
    and full-buffer claims, low occupancy, shallow and deep tasks, a ring of 1024 entries with 3
    slots, tiny slices.  (Shallow tasks (8 free rows) and the small ring get boards up to N=16 only:
    a whole N=19 board would exceed the default task budget, or pass a 1024-entry ring in
    thousands of 3-slice rounds.)"""
    reqs = workloads.poisson_requests(40, lam=1000.0, seed=23)
    if cfg.get("depth_target", 12) < 12 or "slot_tasks" in cfg:
        reqs = [q for q in reqs if q[1] <= 16]
    alone = [nq.instance_count(N, cols) for _, N, cols in reqs]
    with nq.Server(**cfg) as srv:
        tickets = []
        for _, N, cols in reqs:
            for _ in range(60000):
                try:
                    tickets.append(srv.submit(N, cols))
                    break
                except nq.NQError as e:
                    assert e.status == nq.NQ_ERR_BUSY
                    time.sleep(0.001)
        got = [srv.wait(t, timeout_ms=60000)[0] for t in tickets]
    assert got == alone
