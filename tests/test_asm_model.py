"""The hot loop as shipped, run on the CPU (tests/asm_model.py): the six asm blocks of hot_loop in
n_queens_amd/csrc/nq_dfs.hip, taken from the device build's own preprocessing and interpreted for
64-lane waves, give the oracle's per-task counts, totals and node counts, keep every LDS access in
the lane's own slots or the dead frame, restore exec and return within their trip bound.  Planted
mistakes in the instruction text (never in GPU code) must be caught."""
import pytest

import oracle

import asm_model as am

pytestmark = pytest.mark.skipif(am.compiler() is None, reason="needs the HIP compiler to preprocess nq_dfs.hip")

VARIANTS = ("always_wave", "masked_wave", "always_lane", "masked_lane", "always_lane_nodes", "masked_lane_nodes")


@pytest.fixture(scope="module")
def blocks():
    b = am.hot_loop_blocks(am.preprocess())
    assert sorted(b) == sorted(VARIANTS)
    return b


def run(lines, key, N, rows, waves=1, block=64, max_trips=3):
    tasks = oracle.gen_tasks(N, rows)
    return tasks, am.count_tasks(lines, N, tasks, push_always=key.startswith("always"), lane_count="lane" in key,
                                 nodes=key.endswith("nodes"), frames=N - 1 - rows, max_trips=max_trips, waves=waves,
                                 block=block)


@pytest.mark.parametrize("key", VARIANTS)
@pytest.mark.parametrize("N, rows, waves, block", [(6, 1, 1, 64), (8, 2, 2, 128), (9, 3, 1, 64), (10, 2, 2, 128),
                                                    (11, 3, 1, 64)])
def test_shipped_loop_exact(blocks, key, N, rows, waves, block):
    tasks, (per, total, nodes) = run(blocks[key], key, N, rows, waves, block)
    want_sum, want = oracle.count_tasks(N, tasks, per_task=True)
    assert total == want_sum
    if "lane" in key:
        assert (per == want).all()
    if key.endswith("nodes"):
        assert nodes == oracle.count_tasks_nodes(N, tasks)


@pytest.mark.parametrize("key", ["always_wave", "masked_lane_nodes"])
def test_shipped_loop_returns_after_every_trip(blocks, key):
    """A trip bound of 1 makes the loop return after every 16 iterations (the counter folds)."""
    tasks, (per, total, nodes) = run(blocks[key], key, 10, 3, max_trips=1)
    assert total == oracle.count_tasks(10, tasks)
    if key.endswith("nodes"):
        assert nodes == oracle.count_tasks_nodes(10, tasks)


@pytest.mark.parametrize("key", ["always_wave", "masked_wave"])
def test_shipped_loop_whole_board_n12(blocks, key):
    _, (_, total, _) = run(blocks[key], key, 12, 2, waves=2, block=128)
    assert 2 * total == 14200  # A000170


def _plant(lines, old, new, first_only=False):
    out, done = [], False
    for ln in lines:
        if old in ln and not (first_only and done):
            ln = ln.replace(old, new)
            done = True
        out.append(ln)
    assert done, old
    return out


@pytest.mark.parametrize("key, old, new, first_only", [
    # the pop without saturation: a pop from an empty stack reads below the lane's slots
    ("always_wave", "%[stride] clamp", "%[stride]", False),
    # the forced move taken on two candidates instead of one
    ("always_lane", "v_cmpx_eq_u32_e64 %[m1], 1, %[t]", "v_cmpx_eq_u32_e64 %[m1], 2, %[t]", False),
    # the unconditional push advancing two levels
    ("always_wave", "v_min_u32 %[t], 1, v23", "v_min_u32 %[t], 2, v23", False),
    # exec left at the forced lanes after the forced move (the first one only)
    ("masked_lane", "v_bitop3_b32 v23, v21, v22, v20 bitop3:1\ns_mov_b64 exec, %[live]", None, True),
])
def test_planted_mistakes_are_caught(blocks, key, old, new, first_only):
    lines = blocks[key]
    if new is None:
        # drop the exec restore that follows the first forced move
        k = next(i for i in range(len(lines) - 1) if lines[i].startswith("v_bitop3_b32")
                 and lines[i + 1] == "s_mov_b64 exec, %[live]" and any(
                     lines[j].startswith("v_cmpx_eq_u32_e64 %[m1], 1") for j in range(max(0, i - 8), i)))
        planted = lines[:k + 1] + lines[k + 2:]
    else:
        planted = _plant(lines, old, new, first_only)
    N, rows = 8, 2
    tasks = oracle.gen_tasks(N, rows)
    want_sum, want = oracle.count_tasks(N, tasks, per_task=True)
    try:
        per, total, _ = am.count_tasks(planted, N, tasks, push_always=key.startswith("always"),
                                       lane_count="lane" in key, nodes=False, frames=N - 1 - rows)
    except am.LoopError:
        return
    assert total != want_sum or ("lane" in key and not (per == want).all())
