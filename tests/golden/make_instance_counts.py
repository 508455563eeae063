"""Write tests/golden/instance_counts.json: the oracle's count of every instance of the seeded
benchmark workloads, so that batched, per-launch and served GPU counts are checked against the
oracle for every N, not only against another GPU path.

* batch_mix: the 512 instances of n_queens_amd.workloads.batch_mix() (N in [14, 20], 2..6 fixed
  rows, seed 0x6E51);
* poisson: the 400 requests of n_queens_amd.workloads.poisson_requests(400, 1.0) (N in [12, 19],
  0..5 fixed rows, seed 0x5E7E), the set bench.py serves.

An instance's count is the number of completions of its fixed rows, what the reference's
n_queens(N, cur, left, right, sum) returns for the prefix masks (src/n_queens.cpp:8-21), computed
by the oracle's restatement (oracle/nq_oracle.c, count_subtree over OpenMP threads).  Each prefix
is first expanded on the host by enough rows to give the threads many small tasks; the
per-subtask counts are summed back per instance.  A whole board (no fixed row) is counted as
2 x the oracle's count of the reference's mirror-symmetric task list and checked against A000170
(tests/golden/published.json).

    python tests/golden/make_instance_counts.py [--threads 8]

About 15 minutes on 8 cores (the N=19 instances take most of it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from n_queens_amd import workloads  # noqa: E402
from n_queens_amd.workloads import instances_digest  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "instance_counts.json")


def expand(N: int, c: int, l: int, r: int, free_rows_target: int):
    """All placements below (c, l, r) down to `free_rows_target` free rows, as (c, l, r) rows; a
    prefix already at or below the target is returned as it is."""
    last = (1 << N) - 1
    level = [(c, l, r)]
    depth = bin(c).count("1")
    while depth < N and N - depth > free_rows_target:
        nxt = []
        for c0, l0, r0 in level:
            v = last & ~(c0 | l0 | r0)
            while v:
                p = v & -v
                v ^= p
                nxt.append((c0 | p, ((l0 | p) << 1) & 0xFFFFFFFF, (r0 | p) >> 1))
        level = nxt
        depth += 1
    return level


def count_instances(insts, threads: int, log) -> list[int]:
    counts = np.zeros(len(insts), dtype=np.uint64)
    by_n: dict[int, list[tuple[int, tuple[int, int, int]]]] = {}
    for i, (N, cols) in enumerate(insts):
        enc = oracle.encode_prefix(N, cols)
        if enc is None:
            raise ValueError(f"instance {i} is not a valid prefix: {N} {cols}")
        if not cols:
            continue  # whole boards below
        # Subtasks with at most 12 (N < 18) or 14 free rows: many small tasks for the threads.
        for sub in expand(N, *enc, free_rows_target=12 if N < 18 else 14):
            by_n.setdefault(N, []).append((i, sub))
    for N in sorted(by_n):
        idx = np.array([i for i, _ in by_n[N]], dtype=np.int64)
        tasks = np.array([s for _, s in by_n[N]], dtype=np.uint32)
        t0 = time.perf_counter()
        _, per = oracle.count_tasks(N, tasks, threads=threads, per_task=True)
        np.add.at(counts, idx, per)
        log(f"N={N}: {len(tasks)} subtasks of {len(set(idx.tolist()))} instances in {time.perf_counter() - t0:.1f} s")
    # Whole boards: the oracle's generator (the reference's mirror-symmetric task list,
    # src/n_queens.cpp:74-113) and 2 x the sum of its per-task counts (src/n_queens_cuda.cu:181-183).
    whole = {}
    for i, (N, cols) in enumerate(insts):
        if not cols:
            if N not in whole:
                whole[N] = 2 * oracle.count_tasks(N, oracle.gen_tasks(N, min(4, max(1, N - 2))), threads=threads)
            counts[i] = whole[N]
    return [int(x) for x in counts]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)  # noqa: E731
    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        a000170 = {int(k): int(v) for k, v in json.load(f)["solutions"].items()}
    mix = workloads.batch_mix()
    reqs = [(N, cols) for _, N, cols in workloads.poisson_requests(400, lam=1.0)]
    out = {"_about": __doc__.split("\n\n")[0].replace("\n", " "),
           "batch_mix": {"workload": "workloads.batch_mix(): 512 instances, N in [14, 20], 2-6 fixed rows, "
                                     "seed 0x6E51", "digest": instances_digest(mix)},
           "poisson": {"workload": "workloads.poisson_requests(400, 1.0): N in [12, 19], 0-5 fixed rows, "
                                   "seed 0x5E7E", "digest": instances_digest(reqs)}}
    for key, insts in (("batch_mix", mix), ("poisson", reqs)):
        log(f"{key}: {len(insts)} instances")
        counts = count_instances(insts, a.threads, log)
        for (N, cols), c in zip(insts, counts):
            if not cols and c != a000170[N]:
                raise SystemExit(f"whole board N={N}: {c} != A000170 {a000170[N]}")
        out[key]["counts"] = counts
        out[key]["total"] = sum(counts)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")
    log(f"wrote {OUT}")


if __name__ == "__main__":
    main()
