"""GPU parity of the device split done in registers (split_kernel): each wave claims whole tasks
and walks the top `split` rows of their subtrees, handing the roots to its lanes' DFS.  Counts
are integers and must match bit for bit."""
import numpy as np
import pytest

import n_queens_amd as nq
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,rows", [(12, 2), (14, 3), (16, 4), (17, 5), (18, 6)])
def test_split_kernel_totals(N, rows, published):
    """Every split depth 1..8 (scalar-register walk up to 4, frames in lanes beyond), both push forms, narrow and wide workgroups, single and chunked
    task claims: the same total as A000170."""
    t = nq.generate(N, rows)
    for split in range(1, 9):
        for push, block, chunk in ((nq.NQ_PUSH_MASKED, 64, 1), (nq.NQ_PUSH_ALWAYS, 256, 5), (0, 0, 0)):
            r = nq.count(N, t, cfg=nq.Config(split=split, push=push, block=block, chunk=chunk))
            assert 2 * r.sum == published["solutions"][N], (N, rows, split, push, block, chunk)


def test_split_kernel_mixed_depths_vs_oracle():
    """A list with tasks of every depth (full boards, last-row and dead prefixes, the empty
    board), the stack sized for the shallowest: the walk counts the deep ones itself."""
    N = 12
    rng = np.random.default_rng(12)
    parts = [oracle.gen_tasks(N, d) for d in range(1, N + 1)]
    t = np.concatenate([p[rng.choice(len(p), min(len(p), 50), replace=False)] for p in parts if len(p)])
    t = np.concatenate([t, np.zeros((1, 3), np.uint32), np.array([[0b11, 0b1100, 0]], np.uint32)])
    rng.shuffle(t)
    want = oracle.count_tasks(N, t)
    for split in range(0, 9):
        assert nq.count(N, t, cfg=nq.Config(split=split)).sum == want, split


def test_split_per_task_counts_and_nodes():
    """Per-task counts and node counts go through the split too (split_kernel adds each root's
    count to its task; its node count is the walk's placements plus each root's expansions less
    its boards): every task's count and the node count equal the oracle's at every split depth;
    per-task and node counts asked for together take one lane per task (dfs_kernel)."""
    N, rows = 14, 3
    t = nq.generate(N, rows)
    _, want = oracle.count_tasks(N, t, per_task=True)
    want_nodes = oracle.count_tasks_nodes(N, t)
    for split in range(1, 9):
        for push in (nq.NQ_PUSH_MASKED, nq.NQ_PUSH_ALWAYS):
            r = nq.count(N, t, cfg=nq.Config(split=split, push=push), per_task=True)
            assert np.array_equal(r.per_task, want) and r.sum == int(want.sum()), (split, push)
            r = nq.count(N, t, cfg=nq.Config(split=split, push=push), nodes=True)
            assert r.nodes == want_nodes and r.sum == int(want.sum()), (split, push, r.nodes, want_nodes)
    r = nq.count(N, t, cfg=nq.Config(split=3), per_task=True, nodes=True)
    assert np.array_equal(r.per_task, want)
    assert r.nodes == want_nodes


def test_reference_vectors_per_task_through_the_split(ref_vectors):
    """The reference's own per-task counts (its CPU build, tests/golden/ref_vectors.npz), counted
    with the device split walking 1, 2, 3 and 5 rows below each task."""
    for N, rows, total in ref_vectors["totals"]:
        N, rows, total = int(N), int(rows), int(total)
        tasks, counts = ref_vectors[f"N{N}_r{rows}_tasks"], ref_vectors[f"N{N}_r{rows}_counts"]
        for split in (1, 2, 3, 5):
            if rows + split > N - 2:
                continue
            r = nq.count(N, tasks, cfg=nq.Config(split=split), per_task=True)
            assert np.array_equal(r.per_task, counts) and 2 * r.sum == total, (N, rows, split)


def test_n21_per_task_counts(published):
    """N=21, 6 rows: per-task counts through the automatic split sum to the total, and a seeded
    sample of tasks matches the oracle's count of each."""
    N, rows = 21, 6
    t = nq.generate(N, rows)
    r = nq.count(N, t, per_task=True)
    assert 2 * int(r.per_task.sum()) == 2 * r.sum == published["solutions"][N]
    rng = np.random.default_rng(21)
    idx = np.sort(rng.choice(len(t), size=48, replace=False))
    _, want = oracle.count_tasks(N, t[idx], per_task=True)
    assert np.array_equal(r.per_task[idx], want)


def test_split_above_the_limit_is_an_argument_error():
    with pytest.raises(nq.NQError) as e:
        nq.count(12, nq.generate(12, 2), cfg=nq.Config(split=9))
    assert e.value.status == nq.NQ_ERR_ARG


@pytest.mark.parametrize("split", [2, 3, 4])
def test_n20_split_vs_published(split, published):
    total, st = nq.solve(20, 6, nq.Config(split=split))
    assert st["ntasks"] == 2967560
    assert total == published["solutions"][20], split


def test_count_device_split_is_captured_in_a_graph(published):
    """nq_count_device with a split enqueues kernels only (no host synchronisation, no memset),
    so it can be captured in a graph: every replay, with eager work between replays, gives the
    eager count."""
    import torch
    N, rows = 17, 5
    tasks = nq.generate(N, rows)
    dev = torch.device("cuda:0")
    d_tasks = torch.from_numpy(tasks.view(np.int32).reshape(-1)).to(dev)
    d_res = torch.zeros(3, dtype=torch.int64, device=dev)
    cfg = nq.Config(split=2, stack_depth=N - 1 - rows)
    s = torch.cuda.Stream(device=dev)
    # Eager count (also initialises the library's device state outside the capture).
    nq.count_device(N, d_tasks.data_ptr(), len(tasks), d_res.data_ptr(), cfg=cfg, stream=s.cuda_stream)
    s.synchronize()
    eager = d_res.cpu().numpy().view(np.uint64).copy()
    assert 2 * int(eager[0]) == published["solutions"][N] and int(eager[2]) == 0
    g = torch.cuda.CUDAGraph()
    d_res.zero_()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        nq.count_device(N, d_tasks.data_ptr(), len(tasks), d_res.data_ptr(), cfg=cfg,
                        stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i in range(4):
        d_res.zero_()
        junk = torch.arange(1 << 20, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = d_res.cpu().numpy().view(np.uint64)
        assert np.array_equal(got, eager), (i, got, eager)
        assert int(junk) == (1 << 20) * ((1 << 20) - 1) // 2


def test_count_device_per_task_split_is_captured_in_a_graph():
    """Per-task counts through the split, captured in a graph: the per-task array is cleared by a
    kernel (no memset node) before split_kernel adds to it, so every replay, with eager work and a
    poisoned array between replays, gives the eager per-task counts."""
    import torch
    N, rows = 16, 4
    tasks = nq.generate(N, rows)
    dev = torch.device("cuda:0")
    d_tasks = torch.from_numpy(tasks.view(np.int32).reshape(-1)).to(dev)
    d_res = torch.zeros(3, dtype=torch.int64, device=dev)
    d_pt = torch.zeros(len(tasks), dtype=torch.int64, device=dev)
    cfg = nq.Config(split=3, stack_depth=N - 1 - rows)
    s = torch.cuda.Stream(device=dev)
    nq.count_device(N, d_tasks.data_ptr(), len(tasks), d_res.data_ptr(), d_per_task=d_pt.data_ptr(), cfg=cfg,
                    stream=s.cuda_stream)
    s.synchronize()
    eager = d_pt.cpu().numpy().view(np.uint64).copy()
    assert np.array_equal(eager, oracle.count_tasks(N, tasks, per_task=True)[1])
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        nq.count_device(N, d_tasks.data_ptr(), len(tasks), d_res.data_ptr(), d_per_task=d_pt.data_ptr(), cfg=cfg,
                        stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i in range(4):
        d_pt.fill_(7)  # the replay must clear what a previous use left
        junk = torch.arange(1 << 20, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(d_pt.cpu().numpy().view(np.uint64), eager), i
        assert int(junk) == (1 << 20) * ((1 << 20) - 1) // 2


def test_reference_single_gpu_setting_and_automatic_splits(published, monkeypatch):
    """The reference's advice for one GPU (README: N <= 22 with 5 rows pre-placed under
    configuration3) through the CLI at N=20, and N=21 with 5 and 4 rows, where the automatic split
    walks 4 and 5 rows so that the DFS keeps 11 LDS frames (tuner cache off: built-in rule)."""
    import subprocess
    monkeypatch.setenv("NQ_TUNE_CACHE", "")
    p = subprocess.run([nq.CLI_PATH, "20", "5", "--config", "3"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and "cuda 20 queens result 39029188884, calc time:" in p.stdout, p.stdout + p.stderr
    for rows, split in ((5, 4), (4, 5)):
        cfg = nq.solve_config(21, rows)
        assert cfg.split == split and 21 - 1 - (rows + split) == 11, (rows, cfg)
        total, _ = nq.solve(21, rows)
        assert total == published["solutions"][21], rows
