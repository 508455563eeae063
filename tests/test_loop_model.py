"""The DFS loop's control logic on the CPU (tests/loop_model.c): the pop onto the dead frame at
LDS address 0 stays within each lane's own stack slots, every finished lane parks there, and
per-task counts and node counts are exact with the forced move (a child with one candidate has
it placed in the same iteration), for N=4..13, workgroups of 64..256 lanes and both push forms.
Four planted mistakes (a pop without saturation, a dead frame that looks like a full board, a
forced move from a child with no candidate, a count taken before the forced move) must make the
model fail."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "loop_model.c")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


def _build(tmp_path, *defines):
    exe = tmp_path / ("loop_model" + "".join(d.lower() for d in defines))
    cmd = ["gcc", "-O2", "-std=c11", "-o", str(exe), SRC] + [f"-D{d}" for d in defines]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return str(exe)


def test_dead_frame_loop_in_bounds_and_exact(tmp_path):
    out = subprocess.run([_build(tmp_path), "13"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = [s for s in out.stdout.splitlines() if s.endswith(" ok")]
    # N=4..13, rows 1..4 (rows < N-1), 4 workgroup sizes, 2 push forms.
    cases = sum(1 for N in range(4, 14) for rows in range(1, 5) if rows < N - 1) * 4 * 2
    assert len(lines) == cases
    # The N=13 whole-board totals from 1..4 pre-placed rows are all 73,712 (A000170).
    assert all(s.split("sum=")[1].startswith("73712 ") for s in lines if s.startswith("N=13 "))


@pytest.mark.parametrize("plant, message", [("PLANT_NO_CLAMP", "bad read"),
                                            ("PLANT_FULL_DEAD", "expected"),
                                            ("PLANT_FORCED_DEAD", "two boards in one iteration"),
                                            ("PLANT_COUNT_FIRST", "expected")])
def test_planted_mistakes_are_caught(tmp_path, plant, message):
    out = subprocess.run([_build(tmp_path, plant), "8"], capture_output=True, text=True, timeout=300)
    assert out.returncode != 0
    assert message in out.stderr
