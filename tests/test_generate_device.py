"""nq_generate_device: the reference's task list (src/n_queens.cpp:74-113) generated on the
device.  Word for word and in order it equals the host generator (nq_generate, itself pinned to
the reference's own build by tests/golden/ref_vectors.npz), its counts equal the reference logs'
task counts up to N=27 rows 7, it writes nothing past the tasks it returns, and a list generated
on the device and counted there gives A000170."""
import numpy as np
import pytest

import n_queens_amd as nq

pytestmark = pytest.mark.gpu

SENTINEL = -0x2A2A2A2B  # 0xD5D5D5D5 as int32


def _device_list(N, rows, extra=64):
    import torch

    n = nq.generate_device(N, rows)
    buf = torch.full((3 * (n + extra),), SENTINEL, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    got = nq.generate_device(N, rows, buf.data_ptr(), n)
    assert got == n
    host = buf.cpu().numpy()
    assert (host[3 * n:] == SENTINEL).all(), "written past the returned tasks"
    return host[:3 * n].view(np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("N, rows", [(4, 1), (4, 2), (5, 2), (6, 6), (7, 7), (8, 3), (9, 4), (11, 1), (13, 5),
                                     (16, 5), (17, 5), (20, 6), (21, 6)])
def test_device_list_equals_host_generator(N, rows):
    print(f"generate_device N={N} rows={rows}", flush=True)
    host = nq.generate(N, rows)
    dev = _device_list(N, rows)
    assert dev.shape == host.shape
    assert (dev == host).all()


def test_counts_equal_the_reference_logs(published):
    """Every 'Use ... to generate K subproblems!' line of the reference logs, N=20..27."""
    for N, rows, k, where in published["task_counts"]:
        print(f"generate_device count N={N} rows={rows} ({where})", flush=True)
        assert nq.generate_device(N, rows) == k, where


def test_cap_too_small_writes_nothing():
    import torch

    n = nq.generate_device(13, 5)
    buf = torch.full((3 * n,), SENTINEL, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    assert nq.generate_device(13, 5, buf.data_ptr(), n - 1) == n
    assert (buf.cpu().numpy() == SENTINEL).all()


@pytest.mark.parametrize("N, rows", [(16, 5), (17, 4)])
def test_generated_and_counted_on_the_device(N, rows, published):
    """Generation and count without the list crossing PCIe: 2 x sum = A000170(N)."""
    import torch

    n = nq.generate_device(N, rows)
    d_tasks = torch.empty(3 * n, dtype=torch.int32, device="cuda:0")
    d_res = torch.zeros(3, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    assert nq.generate_device(N, rows, d_tasks.data_ptr(), n) == n
    nq.count_device(N, d_tasks.data_ptr(), n, d_res.data_ptr())
    torch.cuda.synchronize()
    res = d_res.cpu().numpy()
    assert res[2] == 0
    assert 2 * int(res[0]) == published["solutions"][N]
