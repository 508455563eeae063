"""Launch-configuration autotuner for one (N, rows) on the current GPU.

Times nq_count (device split + DFS, HIP events around both), checks that every configuration
gives the same total, and caches the fastest per (device, N, rows) in a JSON file.  The counts
never depend on the configuration; only the time does.

* Phase 1: preplace depth (device-split rows on top of `rows`), workgroup size, claim chunk and
  push form, at the most resident waves the LDS stack allows.
* Phase 2 (``--occupancy``): at the phase-1 winner, resident waves per CU (4..16) and LDS stack
  depth (the minimum, +2, +4 frames), with the waves per CU each launch actually had.

    python -m n_queens_amd.tune 21               # rows 6, full sweep, cache the winner
    python -m n_queens_amd.tune 19 --rows 5 --quick
    python -m n_queens_amd.tune 20 --quick --occupancy

Library use: ``tuned_config(N, rows)`` returns the cached Config (or None) for ``solve``.
"""
from __future__ import annotations

import argparse
import dataclasses
import itertools
import json
import os
import sys

from . import NQ_PUSH_ALWAYS, NQ_PUSH_MASKED, Config, count, device_info, generate

CACHE_PATH = os.environ.get("NQ_TUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "n_queens_amd", "tune.json"))


def candidates(N: int, rows: int, quick: bool = False) -> list[Config]:
    """Configurations to time: split rows such that the DFS still has two rows to place."""
    splits = [s for s in range(4) if rows + s + 2 < N] or [0]
    blocks = (64, 128) if quick else (64, 128, 256)
    chunks = (16,) if quick else (4, 16, 64)
    pushes = (0,) if quick else (NQ_PUSH_MASKED, NQ_PUSH_ALWAYS)
    return [Config(block=b, chunk=c, split=s, push=p)
            for s, b, c, p in itertools.product(splits, blocks, chunks, pushes)]


def occupancy_candidates(N: int, rows: int, win: Config) -> list[Config]:
    """Phase 2: waves per CU and stack depth around the phase-1 winner (stack_depth is given for
    the tasks as listed, i.e. before the device split)."""
    base = N - 1 - rows
    out = [dataclasses.replace(win, waves_per_cu=w) for w in (4, 8, 12, 16)]
    out += [dataclasses.replace(win, stack_depth=base + extra) for extra in (0, 2, 4) if base + extra <= 32]
    return out


def _key(N: int, rows: int) -> str:
    name = device_info(0)["name"]
    return f"{name}|N={N}|rows={rows}"


def load_cache(path: str = CACHE_PATH) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tuned_config(N: int, rows: int, path: str = CACHE_PATH) -> Config | None:
    """The cached winner for (N, rows) on this device, or None."""
    entry = load_cache(path).get(_key(N, rows))
    return Config(**entry["config"]) if entry else None


def _time(N: int, tasks, cfgs, reps: int, total: list, log, label: str) -> list[dict]:
    out = []
    for cfg in cfgs:
        best, waves = None, 0
        for _ in range(reps):
            r = count(N, tasks, cfg=cfg)
            if total[0] is None:
                total[0] = r.sum
            if r.sum != total[0]:
                raise RuntimeError(f"configuration {cfg} counted {r.sum}, others {total[0]}")
            if best is None or r.kernel_ms < best:
                best, waves = r.kernel_ms, r.waves
        out.append({"config": dataclasses.asdict(cfg), "kernel_ms": best, "waves": waves})
        log(f"{label} {cfg}: {best:.1f} ms, {waves} waves")
    return out


def tune(N: int, rows: int, quick: bool = False, reps: int = 1, path: str | None = CACHE_PATH,
         log=print, occupancy: bool = False) -> tuple[Config, list[dict]]:
    """Time every candidate (best of `reps`), return the fastest and all timings (phase 2's
    after phase 1's when `occupancy`); store the winner in the cache at `path` (None: do not
    store).  Raises if two configurations disagree on the total."""
    tasks = generate(N, rows)
    total = [None]
    label = f"N={N} rows={rows}"
    results = _time(N, tasks, candidates(N, rows, quick), reps, total, log, label)
    results.sort(key=lambda x: x["kernel_ms"])
    win = Config(**results[0]["config"])
    if occupancy:
        occ = _time(N, tasks, occupancy_candidates(N, rows, win), reps, total, log, label + " occupancy")
        results = results + occ
        best = min(results, key=lambda x: x["kernel_ms"])
        if best["kernel_ms"] < results[0]["kernel_ms"]:
            win = Config(**best["config"])
        results.sort(key=lambda x: x["kernel_ms"])
    total = total[0]
    if path:
        cache = load_cache(path)
        cache[_key(N, rows)] = {"config": results[0]["config"], "kernel_ms": results[0]["kernel_ms"],
                                "solutions": 2 * total, "candidates": len(results)}
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return win, results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("N", type=int)
    ap.add_argument("--rows", type=int, default=6)
    ap.add_argument("--quick", action="store_true", help="split x block only (chunk 16, automatic push)")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--cache", default=CACHE_PATH, help="JSON cache file ('' = do not store)")
    ap.add_argument("--occupancy", action="store_true", help="phase 2: waves/CU and stack depth at the winner")
    a = ap.parse_args(argv)
    win, res = tune(a.N, a.rows, a.quick, a.reps, a.cache or None, log=lambda s: print(s, flush=True),
                    occupancy=a.occupancy)
    print(json.dumps({"N": a.N, "rows": a.rows, "best": dataclasses.asdict(win),
                      "kernel_ms": res[0]["kernel_ms"], "runner_up_ms": res[1]["kernel_ms"] if len(res) > 1 else None}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
