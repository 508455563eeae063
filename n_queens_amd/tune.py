"""Launch autotuner for whole-board counts of size N on the current GPU.

Times the DFS kernel (HIP events, best of ``reps``) on the whole-board task list, checks that
every configuration gives the same total, and writes the fastest to the cache that the library
reads: ``nq_solve`` / ``cuda_n_queens`` / the ``n_queens`` CLI use the entry for N on this device
type whenever they are not given a configuration (``nq_solve_config`` reports it).  Counts never
depend on the configuration; only the time does.  The reference fixes its launch at compile time
(src/macro.h:25-35: block and stack size presets; src/n_queens_cuda.cu:164-167: grid 1024).

Search (coordinate descent, each phase starting from the best so far):

1. depth: the preplace depth of the claimed task list (5..7 by default) x the rows the kernel
   walks below each task in registers (split 0..4), keeping at least three rows for the DFS;
2. launch: workgroup size 64/128/256 x push form (exec-masked / unconditional);
3. claims: chunk 1, 4, 16, 64 (split > 0: the least a wave claims at a time; split 0: the
   tasks a wave hands to its lanes per cursor grab);
4. occupancy: resident waves per CU capped at 4, 8, 12, 16, and one or two LDS frames beyond the
   minimum the tasks need.

``--quick`` runs phase 1 over preplace depths 5..6 and splits 1..3 only.

    python -m n_queens_amd.tune 21                 # full search, cache the winner
    python -m n_queens_amd.tune 20 --quick --cache /tmp/t.txt

Cache: ``$NQ_TUNE_CACHE``, else ``~/.cache/n_queens_amd/tune.txt``; one line per entry
(``<arch> N=<N> preplace=<d> split=<s> block=<b> chunk=<c> push=<p> waves_per_cu=<w>
stack_extra=<e> ms=<t>``), the last line for (arch, N) winning (n_queens_amd/csrc/nq_tune.cpp).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys

from . import NQ_PUSH_ALWAYS, NQ_PUSH_MASKED, Config, count, device_info, generate

CACHE_PATH = os.environ.get("NQ_TUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "n_queens_amd", "tune.txt"))
FIELDS = ("preplace", "split", "block", "chunk", "push", "waves_per_cu", "stack_extra")


@dataclasses.dataclass(frozen=True)
class Point:
    """One candidate: the cache entry's fields."""
    preplace: int
    split: int
    block: int = 0
    chunk: int = 0
    push: int = 0
    waves_per_cu: int = 0
    stack_extra: int = 0

    def config(self, N: int) -> Config:
        """The nq_count configuration timing this point on the depth-`preplace` task list."""
        return Config(block=self.block, chunk=self.chunk, push=self.push, waves_per_cu=self.waves_per_cu,
                      split=self.split, stack_depth=max(1, N - 1 - self.preplace + self.stack_extra))


def depth_points(N: int, preplace=(5, 6, 7), splits=range(0, 5)) -> list[Point]:
    """Phase 1: preplace depth x split, the DFS keeping at least three rows (preplace + split
    <= N - 3), preplace at least 2 (the symmetric generator's restricted rows)."""
    out = [Point(g, s) for g in preplace for s in splits if 2 <= g and g + s <= N - 3]
    if not out:
        g = max(2, min(N - 3, 2))
        out = [Point(g, 0)]
    return out


def launch_points(p: Point) -> list[Point]:
    return [dataclasses.replace(p, block=b, push=q) for b in (64, 128, 256) for q in (NQ_PUSH_MASKED, NQ_PUSH_ALWAYS)]


def chunk_points(p: Point) -> list[Point]:
    return [dataclasses.replace(p, chunk=c) for c in (1, 4, 16, 64)]


def occupancy_points(p: Point) -> list[Point]:
    return ([dataclasses.replace(p, waves_per_cu=w) for w in (4, 8, 12, 16)] +
            [dataclasses.replace(p, stack_extra=e) for e in (1, 2)])


def arch() -> str:
    """The device's gcnArchName, the cache key the library matches (nq_device_info's name is
    "<marketing name> (<gcnArchName>)")."""
    name = device_info(0)["name"]
    return name[name.rfind("(") + 1:-1] if name.endswith(")") else name.split()[-1]


def load_cache(path: str = CACHE_PATH) -> dict:
    """{(arch, N): fields} from the cache file (the last line per key wins)."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].split()
                if len(line) < 2:
                    continue
                kv = dict(x.split("=", 1) for x in line[1:] if "=" in x)
                try:
                    N = int(kv["N"])
                    out[(line[0], N)] = {k: (float(kv[k]) if k == "ms" else int(kv[k]))
                                         for k in (*FIELDS, "ms") if k in kv}
                except (KeyError, ValueError):
                    continue
    except OSError:
        pass
    return out


def tuned_point(N: int, path: str = CACHE_PATH) -> Point | None:
    """The cached winner for N on this device type, or None."""
    e = load_cache(path).get((arch(), N))
    return Point(**{k: e[k] for k in FIELDS}) if e and all(k in e for k in FIELDS) else None


def store(N: int, p: Point, ms: float, path: str = CACHE_PATH) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    line = " ".join([arch(), f"N={N}", *(f"{k}={getattr(p, k)}" for k in FIELDS), f"ms={ms:.3f}"])
    with open(path, "a") as f:
        f.write(line + "\n")


def tune(N: int, quick: bool = False, reps: int = 1, path: str | None = CACHE_PATH, log=print,
         preplace=(5, 6, 7)) -> tuple[Point, list[dict]]:
    """Run the search; return the winner and every timing.  Raises if two configurations
    disagree on the total.  Stores the winner in the cache at `path` (None: do not store)."""
    lists = {}
    total = [None]
    results: list[dict] = []
    seen: dict[Point, float] = {}

    def time_points(points, phase):
        for p in points:
            if p in seen:
                continue
            if p.preplace not in lists:
                lists[p.preplace] = generate(N, p.preplace)
            best = None
            for _ in range(reps):
                r = count(N, lists[p.preplace], cfg=p.config(N))
                if total[0] is None:
                    total[0] = r.sum
                if r.sum != total[0]:
                    raise RuntimeError(f"{p} counted {r.sum}, other configurations {total[0]}")
                best = r.kernel_ms if best is None else min(best, r.kernel_ms)
            seen[p] = best
            results.append({"phase": phase, **dataclasses.asdict(p), "kernel_ms": best, "waves": r.waves})
            log(f"N={N} {phase:9s} {p}: {best:.2f} ms, {r.waves} waves")
        return min(seen, key=seen.get)

    if quick:
        win = time_points(depth_points(N, tuple(g for g in preplace if g <= 6)[:2] or preplace, range(1, 4)),
                          "depth")
    else:
        win = time_points(depth_points(N, preplace), "depth")
        win = time_points(launch_points(win), "launch")
        win = time_points(chunk_points(win), "claims")
        win = time_points(occupancy_points(win), "occupancy")
    if path:
        store(N, win, seen[win], path)
    return win, results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("N", type=int)
    ap.add_argument("--quick", action="store_true", help="phase 1 only, preplace 5..6, split 1..3")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--preplace", type=int, nargs="+", default=[5, 6, 7], help="preplace depths of phase 1")
    ap.add_argument("--cache", default=CACHE_PATH, help="cache file ('' = do not store)")
    ap.add_argument("--json", help="write every timing to this JSON file")
    a = ap.parse_args(argv)
    win, res = tune(a.N, a.quick, a.reps, a.cache or None, log=lambda s: print(s, flush=True),
                    preplace=tuple(a.preplace))
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"N": a.N, "arch": arch(), "best": dataclasses.asdict(win), "timings": res}, f, indent=1)
    best = min(r["kernel_ms"] for r in res)
    print(json.dumps({"N": a.N, "best": dataclasses.asdict(win), "kernel_ms": best, "candidates": len(res)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
