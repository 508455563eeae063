"""Write tests/golden/mix_nodes.json: the DFS node total of the 512-instance batched mix
(n_queens_amd.workloads.batch_mix(), N in [14, 20], 2..6 fixed rows, seed 0x6E51), so that
bench.py can report the batched kernel's node rate beside the whole-board kernel's.

A node is one queen placed at depths k+1..N-1 below an instance's k fixed rows (the reference
kernel's loop iteration, src/n_queens_cuda.cu:39-72), counted by the oracle's restatement
(oracle/nq_oracle.c, oracle_count_nodes) one instance per thread.

    python tests/golden/make_mix_nodes.py [--threads 8]

About 6 minutes on 8 cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from n_queens_amd import workloads  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mix_nodes.json")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    mix = workloads.batch_mix()

    def nodes(inst):
        N, cols = inst
        enc = oracle.encode_prefix(N, cols)
        if enc is None or not cols:
            raise ValueError(f"the mix holds prefixed instances only: {N} {cols}")
        return oracle.count_nodes(N, *enc)

    # ctypes releases the GIL around each oracle call, so the threads count in parallel.
    with ThreadPoolExecutor(a.threads) as ex:
        per = list(ex.map(nodes, mix))
    out = {"workload": "workloads.batch_mix(): 512 instances, N in [14, 20], 2-6 fixed rows, seed 0x6E51",
           "digest": workloads.instances_digest(mix), "nodes": per, "total": sum(per)}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")
    print(f"wrote {OUT}: {sum(per)} nodes", flush=True)


if __name__ == "__main__":
    main()
