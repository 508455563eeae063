"""Write tests/golden/ref_vectors.npz from the reference's own CPU code.

Runs only where the reference checkout exists (it builds oracle/_ref/libnq_ref.so from the
reference's src/n_queens.cpp + src/utils.cpp with oracle/Makefile).  The committed output is
data only: for each (N, rows) case the task list the reference generates (partial_n_queens +
partial_n_queens_for_odd, src/n_queens.cpp:74-113, in src/n_queens_cuda.cu:91-95 order), the
reference's per-task counts (n_queens, src/n_queens.cpp:8-21) and its serial_n_queens total
(src/n_queens.cpp:115-137).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")

# (N, rows): every legal rows for small boards, plus the N=13 rows=5 parity case.
CASES = [(N, rows) for N in range(4, 12) for rows in range(2, min(N - 2, 5) + 1)] + [(12, 4), (13, 5)]


def main() -> None:
    if oracle.ref() is None:
        oracle.build()
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libnq_ref.so could not be built (no reference checkout)")
    arrays = {}
    totals = []
    for N, rows in CASES:
        t = oracle.ref_gen_tasks(N, rows)
        counts = np.array([R.ref_count_subtree(N, int(c), int(l), int(r)) for c, l, r in t], dtype=np.uint64)
        total = int(R.ref_serial_n_queens(N, rows))
        arrays[f"N{N}_r{rows}_tasks"] = t.astype(np.uint32)
        arrays[f"N{N}_r{rows}_counts"] = counts
        totals.append((N, rows, total))
    arrays["totals"] = np.array(totals, dtype=np.uint64)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(CASES)} cases, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
