"""Benchmark: N-Queens counting on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 21] [--rows 6]

A step is one full count of the N=21 board with 6 rows pre-placed (4,470,757 prefix tasks,
the reference's single-GPU configuration, README.md:24,28), with the task list already
resident in HBM.  The value is DFS nodes per second over the timed steps (a node is one queen
placed at depths rows+1..N-1).  Also reported:
  * n21_wall_s: the whole `n_queens 21 6` CLI call (generation, device init, copies, kernel) —
    the reference's own measure (src/main.cpp:49-54; 21.242 s on one RTX 5090, README.md:24);
  * roofline: integer-VALU bound of the DFS kernel (11 int32 ops per node), with the measured
    binding resource and the kernel's CU cycles per queen placed by one wave;
  * cpu_baseline: the reference's own OpenMP path (oracle/_ref, parallel_n_queens, which fixes
    32 threads) timed on this host on a bounded sample, and cpu_baseline_port: the oracle's
    restatement on every usable CPU of this host (one thread per CPU), same sample;
  * per_task_counts: the same count with one count per task (the reference kernel's own output,
    partial_sum), timed once, its sum checked against the total;
  * batched_mix and serving: the 512-instance mix and the 400 Poisson requests, every count
    checked against the oracle's (tests/golden/instance_counts.json).
Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "N=21 wall s + DFS nodes/s on 1 GPU; batched-mix instances/s vs per-launch"
OPS_PER_NODE = 11          # p=v&-v (2), v^=p (1), c|=p (1), l=(l|p)<<1 (2), r=(r|p)>>1 (2), v'=last&~(c|l|r) (3)
PEAK_CLOCK_HZ = 2.4e9      # MI355X peak engine clock
LANES_PER_CU = 64          # int32 VALU lanes per CU per clock (4 SIMD x 16)
# HBM bytes per N=21 rows-6 launch of the DFS kernel (split_kernel, 3 rows walked in registers),
# from two rocprofv3 --pmc passes scaled by a 256 MiB copy (tests/probes/copy_probe: FETCH_SIZE
# x2.000, WRITE_SIZE x1.000): profiles/split_kernel_n21_rows6_traffic.md.
TRAFFIC_N21_ROWS6 = round((2 * 63816.6875 + 3429.6875) * 1024)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus() -> tuple[int, str]:
    """Usable CPUs of this process and the CPU model name."""
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return ncpu, model


def cpu_baselines(nodes_fn) -> tuple[dict | None, dict]:
    """Whole N=18 board with 6 rows pre-placed on the host: the reference's OpenMP path
    (src/n_queens.cpp:139-163, omp_set_num_threads(32)) and the oracle's restatement with one
    thread per usable CPU."""
    import oracle
    N, rows = 18, 6
    ncpu, model = host_cpus()
    nodes = nodes_fn(N, rows)
    host = f"host: {ncpu} usable CPUs, {model}"
    ref = None
    R = oracle.ref()
    if R is not None:
        t0 = time.perf_counter()
        total = int(R.ref_parallel_n_queens(N, rows))
        dt = time.perf_counter() - t0
        ref = {"value": nodes / dt, "unit": "nodes/s", "cores": ncpu, "threads": 32, "kind": "reference",
               "sample": f"whole N={N} board, {rows} rows pre-placed: the reference's parallel_n_queens "
                         f"(32 OpenMP threads, as it fixes them) in {dt:.2f} s, {total} solutions, {nodes} "
                         f"DFS nodes; {host}",
               "seconds": dt, "solutions": total}
    t0 = time.perf_counter()
    total = oracle.parallel_n_queens(N, rows, threads=ncpu)
    dt = time.perf_counter() - t0
    port = {"value": nodes / dt, "unit": "nodes/s", "cores": ncpu, "threads": ncpu, "kind": "port",
            "sample": f"whole N={N} board, {rows} rows pre-placed: the oracle's parallel_n_queens restatement "
                      f"(one OpenMP thread per usable CPU) in {dt:.2f} s, {total} solutions; {host}",
            "seconds": dt, "solutions": total}
    return ref, port


def instance_fixture(name: str) -> dict:
    """The oracle's count of every instance of a seeded workload (tests/golden/instance_counts.json)."""
    with open(os.path.join(ROOT, "tests", "golden", "instance_counts.json")) as f:
        return json.load(f)[name]


def batched_mix(nq, log) -> dict:
    """512 seeded instances (N in [14,20], 2-6 fixed rows): one batched launch vs one plain
    launch per instance, both host-in/host-out, counts required identical."""
    from n_queens_amd import workloads
    mix = workloads.batch_mix()
    n = len(mix)
    nq.batch_count(mix[:8])  # warm the batch path
    nq.instance_count(*mix[0])
    t0 = time.perf_counter()
    per = [nq.instance_count(N, cols) for N, cols in mix]
    t_per = time.perf_counter() - t0
    t0 = time.perf_counter()
    counts, st = nq.batch_count(mix)
    t_batch = time.perf_counter() - t0
    plan = nq.BatchPlan(mix)
    plan.run()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        c2, st2 = plan.run()
    t_replay = (time.perf_counter() - t0) / reps
    plan.close()
    same = [int(x) for x in counts] == per and [int(x) for x in c2] == per
    fx = instance_fixture("batch_mix")
    digest = workloads.instances_digest(mix)
    oracle_ok = fx["digest"] == digest and per == fx["counts"]
    # The mix's DFS node total (the oracle's, tests/golden/mix_nodes.json): the batched kernel's
    # node rate, comparable with the whole-board kernel's nodes/s.
    mix_nodes = None
    try:
        with open(os.path.join(ROOT, "tests", "golden", "mix_nodes.json")) as f:
            mn = json.load(f)
        if mn["digest"] == digest:
            mix_nodes = int(mn["total"])
    except (OSError, KeyError, ValueError):
        pass
    out = {"instances": n, "per_launch_inst_per_s": n / t_per, "batch_inst_per_s": n / t_batch,
           "batch_replay_inst_per_s": n / t_replay, "speedup": t_per / t_batch,
           "speedup_replay": t_per / t_replay, "counts_identical": same,
           "counts_match_oracle_fixture": oracle_ok, "per_launch_s": t_per,
           "batch_s": t_batch, "batch_replay_s": t_replay, "batch_kernel_ms": st2["kernel_ms"],
           "ntasks": int(st["ntasks"]), "solutions": int(sum(per)), "nodes": mix_nodes,
           "batch_kernel_nodes_per_s": mix_nodes / (st2["kernel_ms"] / 1e3) if mix_nodes else None,
           "workload": "512 instances, N uniform in [14,20], 2-6 fixed rows, seed 0x6E51"}
    log(f"batched mix: per-launch {t_per:.3f}s  batch {t_batch:.3f}s  replay {t_replay:.3f}s  "
        f"speedup {t_per / t_batch:.2f}x (replay {t_per / t_replay:.2f}x)  identical={same} oracle={oracle_ok}")
    return out


def serving(nq, log, n_req: int = 400, util: float = 0.7) -> dict:
    """Poisson arrivals (seeded) of requests with N in [12,19] and a 0-5 row prefix, at a rate
    set to `util` of the measured service capacity (all requests offered at once); p50/p99
    latency, the same split into whole boards and prefixed requests, slowdown (latency over the
    request's own time counted alone with one launch), solutions/s; every served count checked
    against the same request counted alone and against the oracle's."""
    import threading
    from n_queens_amd import workloads
    reqs0 = workloads.poisson_requests(n_req, lam=1.0)
    nq.instance_count(*reqs0[0][1:])  # warm the one-launch path
    alone, alone_ms = [], []
    for _, N, cols in reqs0:
        t0 = time.perf_counter()
        alone.append(nq.instance_count(N, cols))
        alone_ms.append((time.perf_counter() - t0) * 1e3)
    fx = instance_fixture("poisson")
    oracle_ok = fx["digest"] == workloads.instances_digest([(N, cols) for _, N, cols in reqs0]) and alone == fx["counts"]
    busy = [0]

    def submit(srv, N, cols, deadline_s: float = 60.0):
        # The server refuses work beyond its pending-task budget (NQ_ERR_BUSY): back off and retry,
        # for at most deadline_s (a request that stays refused that long is an error).
        t_end = time.perf_counter() + deadline_s
        while True:
            try:
                return srv.submit(N, cols)
            except nq.NQError as e:
                if e.status != nq.NQ_ERR_BUSY:
                    raise
                if time.perf_counter() > t_end:
                    raise RuntimeError(f"request N={N} {cols} refused as busy for {deadline_s} s") from e
                busy[0] += 1
                time.sleep(0.0005)

    with nq.Server() as srv:
        # capacity: all requests offered at once
        t0 = time.perf_counter()
        tickets = [submit(srv, N, cols) for _, N, cols in reqs0]
        burst = [srv.wait(t, timeout_ms=120000)[0] for t in tickets]
        log(f"serving capacity probe done: {srv.stats()}")
        cap_s = time.perf_counter() - t0
    busy_burst, busy[0] = busy[0], 0
    capacity = n_req / cap_s
    lam = util * capacity
    reqs = [(at / lam, N, cols) for at, N, cols in reqs0]  # same requests, arrival times scaled
    tickets = [None] * n_req
    with nq.Server() as srv:
        t0 = time.perf_counter()

        def feed():
            for i, (at, N, cols) in enumerate(reqs):
                d = at - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                tickets[i] = submit(srv, N, cols)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = [srv.wait(t, timeout_ms=120000) for t in tickets]
        span = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([x["latency_ms"] for _, x in res])
    served = [c for c, _ in res]
    same = served == alone and burst == alone
    slow = lat / np.maximum(np.array(alone_ms), 1e-3)
    whole = np.array([not cols for _, _, cols in reqs0])

    def pct(x):
        return {"n": int(len(x)), "p50_ms": float(np.percentile(x, 50)), "p99_ms": float(np.percentile(x, 99))}
    out = {"requests": n_req, "capacity_req_per_s": capacity, "lambda_req_per_s": lam, "target_util": util,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "solutions_per_s": sum(served) / span, "req_per_s": n_req / span,
           "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
           "slowdown_p50": float(np.percentile(slow, 50)), "slowdown_p99": float(np.percentile(slow, 99)),
           "whole_boards": pct(lat[whole]), "prefixed": pct(lat[~whole]),
           "alone_total_s": sum(alone_ms) / 1e3,
           "alone_p50_ms": float(np.percentile(alone_ms, 50)), "alone_p99_ms": float(np.percentile(alone_ms, 99)),
           "batches": int(st["batches"]), "counts_identical_alone_vs_mix": same,
           "counts_match_oracle_fixture": oracle_ok and served == fx["counts"], "span_s": span,
           "busy_retries": busy[0], "busy_retries_capacity_probe": busy_burst,
           "device_busy_ms": st["device_ms"],
           "workload": "Poisson arrivals, N uniform in [12,19], 0-5 fixed rows, seed 0x5E7E"}
    log(f"serving: capacity {capacity:.1f} req/s, lambda {lam:.1f}/s, p50 {out['p50_ms']:.2f} ms, "
        f"p99 {out['p99_ms']:.2f} ms, {out['solutions_per_s']:.3e} solutions/s, batches {st['batches']}, "
        f"identical={same}")
    return out


def device_generation(nq, log) -> dict:
    """The reference's task list generated on the device (nq_generate_device) into a resident
    device buffer, for N=21 rows 6 and N=27 rows 7 (453,688,251 tasks, 5.4 GB): task counts
    checked against the reference logs, best of 3 synchronous calls beside the reference's host
    generation time (one thread, its 5090 log)."""
    import torch

    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        fx = json.load(f)
    logged = {(n, r): k for n, r, k, _ in fx["task_counts"]}
    out = {}
    for N, rows in ((21, 6), (27, 7)):
        n = nq.generate_device(N, rows)  # the count alone
        buf = torch.empty(3 * n, dtype=torch.int32, device="cuda:0")
        torch.cuda.synchronize()
        nq.generate_device(N, rows, buf.data_ptr(), n)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            got = nq.generate_device(N, rows, buf.data_ptr(), n)
            best = min(best, time.perf_counter() - t0)
        del buf
        torch.cuda.empty_cache()
        ref_ms = fx["generation_ms"][f"{N},{rows}"][0]
        out[f"{N},{rows}"] = {"tasks": got, "matches_log": got == logged.get((N, rows)), "ms": best * 1e3,
                              "reference_host_ms": ref_ms, "bytes_written": 12 * got,
                              "write_GB_per_s": 12 * got / best / 1e9}
        log(f"device generation N={N} rows={rows}: {got} tasks in {best * 1e3:.2f} ms "
            f"(reference host {ref_ms} ms), log count {'ok' if got == logged.get((N, rows)) else 'DIFFERS'}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=21)
    ap.add_argument("--rows", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nodes", action="store_true", help="take the node count from the fixtures (profiling)")
    ap.add_argument("--no-wall", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--no-gen", action="store_true", help="skip the device task-generation figures")
    ap.add_argument("--gen-only", action="store_true")
    ap.add_argument("--split", type=int, default=-1, help="device split rows (-1: nq_solve's choice)")
    ap.add_argument("--block", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU "
              f"path (asked for --gpus {args.gpus}).", file=sys.stderr)
        return 2

    global np
    import numpy as np
    import torch

    import n_queens_amd as nq

    N, rows = args.n, args.rows
    info = nq.device_info(0)
    log(f"device {info}")
    if args.batch_only:
        print(json.dumps({"batched_mix": batched_mix(nq, log)}), flush=True)
        return 0
    if args.serve_only:
        print(json.dumps({"serving": serving(nq, log)}), flush=True)
        return 0
    if args.gen_only:
        print(json.dumps({"device_generation": device_generation(nq, log)}), flush=True)
        return 0
    t0 = time.perf_counter()
    tasks = nq.generate(N, rows)
    gen_s = time.perf_counter() - t0
    ntasks = len(tasks)
    log(f"generated {ntasks} tasks for N={N} rows={rows} in {gen_s*1e3:.1f} ms")

    dev = torch.device("cuda:0")
    d_tasks = torch.from_numpy(tasks.view(np.int32).reshape(-1)).to(dev)
    d_result = torch.zeros(3, dtype=torch.int64, device=dev)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    # nq_solve's configuration: stack_depth bounds the tasks as uploaded (depth `rows`); the
    # library takes the split rows off it for the DFS after the split.
    cfg = nq.solve_config(N, rows)
    if args.split >= 0:
        cfg = nq.config_default(N, rows)
        cfg.split = args.split
    if args.block:
        cfg.block = args.block
    cfg_nodes = nq.config_default(N, rows)
    cfg_nodes.block = cfg.block

    def run(count_nodes: bool, c=None) -> None:
        # The node-counting run counts the unsplit list (nodes at depths rows+1..N-1).
        nq.count_device(N, d_tasks.data_ptr(), ntasks, d_result.data_ptr(),
                        cfg=c or (cfg_nodes if count_nodes else cfg), count_nodes=count_nodes,
                        stream=stream.cuda_stream)

    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        fixtures = json.load(f)
    want_nodes = fixtures.get("dfs_nodes", {}).get(f"{N},{rows}")
    want_total = fixtures["solutions"].get(str(N))
    if args.no_nodes:
        # Profiling runs: the node count and total come from the fixtures (checked by the GPU tests
        # and by every full bench run); the node-counting kernel is not run.
        if want_nodes is None or want_total is None:
            log("ERROR: --no-nodes needs fixture node and solution counts for this (N, rows)")
            return 1
        nodes, sum0 = int(want_nodes), int(want_total) // 2
    else:
        # Node count (node-counting kernel, once, untimed; not a warmup of the timed path).
        d_result.zero_()
        torch.cuda.synchronize()
        run(True)
        stream.synchronize()
        res = d_result.cpu().numpy().view(np.uint64)
        sum0, nodes = int(res[0]), int(res[1])
        log(f"node-counting run: sum={sum0} total={2*sum0} nodes={nodes} flags={int(res[2])}")
        if int(res[2]):
            log("ERROR: device error flags set")
            return 1
        if (want_nodes is not None and nodes != want_nodes) or (rows >= 2 and want_total is not None
                                                                 and 2 * sum0 != want_total):
            log(f"ERROR: nodes {nodes} / total {2 * sum0} differ from the fixtures {want_nodes} / {want_total}")
            return 1
    # Warmup: W untimed steps of exactly the timed work.
    for i in range(args.warmup):
        d_result.zero_()
        run(False)
        stream.synchronize()
        log(f"warmup {i+1} done")

    # Timed steps: events on the launch stream.
    d_result.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        run(False)
        ev[i][1].record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    step_ms = [a.elapsed_time(b) for a, b in ev]
    res = d_result.cpu().numpy().view(np.uint64)
    sums = int(res[0])
    ok = sums == args.steps * sum0 and int(res[2]) == 0
    log(f"timed steps ms={['%.1f' % x for x in step_ms]} wall={wall:.3f}s counts_consistent={ok}")
    if not ok:
        log("ERROR: counts differ between steps")
        return 1
    ms_per_step = sum(step_ms) / len(step_ms)
    nodes_per_s = nodes / (ms_per_step / 1e3)

    per_task = None
    if not args.no_nodes:
        # The reference kernel's own output, one count per task (partial_sum,
        # src/n_queens_cuda.cu:79), through the same split: one run, HIP events on the stream.
        d_per_task = torch.zeros(ntasks, dtype=torch.int64, device=dev)
        d_result.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        nq.count_device(N, d_tasks.data_ptr(), ntasks, d_result.data_ptr(), d_per_task=d_per_task.data_ptr(), cfg=cfg,
                        stream=stream.cuda_stream)
        e1.record(stream)
        stream.synchronize()
        pt = d_per_task.cpu().numpy().view(np.uint64)
        pt_ok = int(pt.sum()) == sum0 == int(d_result.cpu().numpy().view(np.uint64)[0])
        per_task = {"ms": e0.elapsed_time(e1), "tasks": ntasks, "sum_matches_total": pt_ok,
                    "kernel": "split_kernel with per-task accumulation (the reference kernel's partial_sum)"}
        log(f"per-task counts: {per_task['ms']:.1f} ms, sum matches the total: {pt_ok}")
        del d_per_task
        if not pt_ok:
            log("ERROR: per-task counts do not sum to the total")
            return 1
    peak = info["cus"] * LANES_PER_CU * PEAK_CLOCK_HZ / 1e12
    achieved = nodes * OPS_PER_NODE / (ms_per_step / 1e3) / 1e12
    # The DFS launch of a timed step as the library resolves it (the split tasks start at depth
    # rows + split): workgroup size, push form, grid.
    split = max(0, cfg.split)
    launch = nq.resolve_config(N, min(N, rows + split), 1 << 40, nq.Config(
        block=cfg.block, stack_depth=max(1, cfg.stack_depth - split), chunk=cfg.chunk, push=cfg.push,
        waves_per_cu=cfg.waves_per_cu, split=split))

    out = {
        "metric": METRIC, "value": nodes_per_s, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"N={N} whole-board count, {rows} rows pre-placed (configuration2)", "N": N,
                   "rows": rows, "ntasks": ntasks, "block": launch.block, "dfs_stack_frames": launch.stack_depth,
                   "push": "unconditional" if launch.push == nq.NQ_PUSH_ALWAYS else "masked",
                   "waves_per_cu": launch.grid * (launch.block // 64) // info["cus"],
                   "chunk": launch.chunk, "device_split_rows": cfg.split,
                   "parallelism": "single GPU, persistent grid"},
        "solutions": 2 * sum0, "nodes": nodes, "per_task_counts": per_task,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tops/s (int32 lane-ops)",
                     "frac": achieved / peak,
                     "traffic": TRAFFIC_N21_ROWS6 if (N, rows) == (21, 6) else None,
                     "basis": f"{OPS_PER_NODE} int32 ops per DFS node x {nodes} nodes / kernel time; peak = "
                              f"{info['cus']} CUs x 64 lanes x 2.4 GHz",
                     # What binds the loop, from the SQ counters of this launch
                     # (profiles/split_kernel_n21_rows6_sq_counters.md), and the CU cycles one
                     # wave spends per queen placed (nodes and completed boards).
                     "binding": "VALU issue, 0.97 of the wave-instruction peak: 20 VALU, one ds_write_b128 "
                                "push and one ds_read_b128 pop per iteration, 1.37 placements per iteration "
                                "(forced moves)",
                     "cu_cycles_per_wave_placement": (ms_per_step / 1e3) * PEAK_CLOCK_HZ * info["cus"]
                                                     / ((nodes + sum0) / 64.0)},
    }

    if not args.no_wall:
        t0 = time.perf_counter()
        p = subprocess.run([nq.CLI_PATH, str(N), str(rows)], capture_output=True, text=True, timeout=900)
        outer = time.perf_counter() - t0
        m = re.search(r"cuda (\d+) queens result (-?\d+), calc time: \[([\d.]+)ms\]", p.stdout)
        if m:
            out["n21_wall_s" if N == 21 else "wall_s"] = float(m.group(3)) / 1e3
            out["wall_process_s"] = outer
            out["wall_result_ok"] = int(m.group(2)) == 2 * sum0
            out["wall_vs_rtx5090"] = (21.242 / (float(m.group(3)) / 1e3)) if N == 21 else None
            log(f"CLI: {m.group(0)} (process {outer:.2f} s)")
        else:
            log(f"CLI output not parsed: {p.stdout[-500:]} {p.stderr[-500:]}")

    if not args.no_batch:
        out["batched_mix"] = batched_mix(nq, log)
    if not args.no_serve:
        out["serving"] = serving(nq, log)
    if not args.no_gen:
        out["device_generation"] = device_generation(nq, log)

    if not args.no_cpu:
        def nodes_fn(n, r):
            t = nq.generate(n, r)
            return nq.count(n, t, nodes=True).nodes
        ref, port = cpu_baselines(nodes_fn)
        out["cpu_baseline"] = ref or port
        out["cpu_baseline_port"] = port
        # Not measured: the N=21 count's node total over each baseline's rate.
        out["cpu_n21_estimate_s" if N == 21 else "cpu_estimate_s"] = {
            "reference_32_threads": nodes / ref["value"] if ref else None, "port_all_cpus": nodes / port["value"],
            "basis": "estimate: this count's DFS nodes / the sample's nodes per second (not timed)"}
        for b in (ref, port):
            if b:
                log(f"cpu baseline ({b['kind']}, {b['threads']} threads) {b['value']:.3e} nodes/s ({b['seconds']:.1f} s)")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
