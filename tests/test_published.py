"""tests/golden/published.json against the reference's own files: every task count, host
generation time and logged solution count it cites is on the cited line.  Runs where the
reference is mounted (the build container); elsewhere the fixture stands as committed."""
import json
import os
import re

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "logs")), reason="reference not mounted")


@pytest.fixture(scope="module")
def pub():
    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        return json.load(f)


def _line(where):
    path, no = where.rsplit(":", 1)
    with open(os.path.join(REF, path)) as f:
        return f.read().splitlines()[int(no) - 1]


GEN = re.compile(r"Use ([0-9.]+)ms to generate (\d+) subproblems!")


def test_task_counts_and_generation_times_are_on_the_cited_lines(pub):
    for N, rows, k, where in pub["task_counts"]:
        m = GEN.search(_line(where))
        assert m and int(m.group(2)) == k, where
        ms, where2 = pub["generation_ms"][f"{N},{rows}"]
        assert where2 == where and float(m.group(1)) == ms, where


def test_logged_solution_counts_are_on_the_cited_lines(pub):
    for n, where in pub["solutions_source"].items():
        if "-" in n:
            continue
        m = re.search(r"cuda (\d+) queens result (\d+)", _line(where))
        assert m and int(m.group(1)) == int(n) and int(m.group(2)) == int(pub["solutions"][n]), where
