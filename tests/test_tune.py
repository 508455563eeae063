"""The autotuner drives the product: a quick search at N=20 writes the cache entry that
nq_solve_config reports and that nq_solve, cuda_n_queens and the CLI then use, bit-exact."""
import os
import re
import subprocess

import pytest

import n_queens_amd as nq
from n_queens_amd import tune

pytestmark = pytest.mark.gpu

N20 = 39029188884


def test_quick_tune_n20_drives_nq_solve(tmp_path, monkeypatch):
    path = str(tmp_path / "tune.txt")
    win, res = tune.tune(20, quick=True, path=path, log=lambda s: None)
    assert len(res) == len(tune.depth_points(20, (5, 6), range(1, 4)))
    assert tune.tuned_point(20, path) == win
    monkeypatch.setenv("NQ_TUNE_CACHE", path)
    cfg = nq.solve_config(20, 6)
    g = max(6, win.preplace)
    assert (cfg.preplace, cfg.split) == (g if g > 6 else 0, max(0, win.preplace + win.split - g)), (cfg, win)
    total, st = nq.solve(20, 6)
    assert total == N20 and st["ntasks"] == 2967560
    # The CLI (cuda_n_queens) reads the same cache.
    p = subprocess.run([nq.CLI_PATH, "20", "6"], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, NQ_TUNE_CACHE=path))
    assert f"cuda 20 queens result {N20}," in p.stdout, p.stdout + p.stderr
    assert "to generate 2967560 subproblems!" in p.stdout


def test_tuned_entries_apply_and_keep_the_task_count(tmp_path, monkeypatch, published):
    """Hand-written entries: a preplace deeper than rows (generated further on the device), one
    shallower (raised to rows, split shortened to keep the DFS depth), explicit launch fields;
    a malformed line is skipped.  Every count stays exact and the task count reported is the
    reference's for (N, rows)."""
    a = tune.arch()
    path = tmp_path / "t.txt"
    path.write_text(f"{a} N=17 preplace=7 split=2 block=64 chunk=4 push=1 waves_per_cu=8 stack_extra=1\n"
                    f"{a} N=18 preplace=4 split=3 block=256 chunk=1 push=2 waves_per_cu=0 stack_extra=0\n"
                    f"{a} N=16 preplace=x split=2\n")
    monkeypatch.setenv("NQ_TUNE_CACHE", str(path))
    c17 = nq.solve_config(17, 5)
    assert (c17.preplace, c17.split, c17.block, c17.chunk, c17.push, c17.waves_per_cu, c17.stack_depth) == \
        (7, 2, 64, 4, 1, 8, 17 - 1 - 7 + 1)
    total, st = nq.solve(17, 5)
    assert total == published["solutions"][17] and st["ntasks"] == len(nq.generate(17, 5))
    c18 = nq.solve_config(18, 5)
    assert (c18.preplace, c18.split) == (0, 2)
    total, st = nq.solve(18, 5)
    assert total == published["solutions"][18] and st["ntasks"] == len(nq.generate(18, 5))
    assert nq.solve_config(16, 4).preplace == 0  # the malformed entry is ignored
    assert nq.solve(16, 4)[0] == published["solutions"][16]
