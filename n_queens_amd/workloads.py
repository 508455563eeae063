"""Seeded synthetic workloads of the benchmark configurations.

* ``batch_mix``: 512 independent instances, N uniform in [14, 20], each with a random valid
  non-attacking prefix of 2..6 fixed rows (seed 0x6E51).
* ``poisson_requests``: serving requests with N uniform in [12, 19] and a valid prefix of
  0..5 rows, arriving as a Poisson process of rate ``lam`` per second.

A prefix is built row by row, choosing uniformly among the columns that no earlier queen
attacks; when a row has none, the prefix is drawn again.  Prefixes with zero completions are
allowed.
"""
from __future__ import annotations

import hashlib

import numpy as np

BATCH_SEED = 0x6E51


def random_prefix(rng: np.random.Generator, N: int, k: int) -> list[int]:
    while True:
        cols: list[int] = []
        ok = True
        for row in range(k):
            free = [c for c in range(N)
                    if all(c != pc and abs(c - pc) != row - pr for pr, pc in enumerate(cols))]
            if not free:
                ok = False
                break
            cols.append(int(free[rng.integers(len(free))]))
        if ok:
            return cols


def batch_mix(n: int = 512, seed: int = BATCH_SEED, n_range=(14, 20), k_range=(2, 6)) -> list[tuple[int, list[int]]]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.integers(n_range[0], n_range[1] + 1))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        out.append((N, random_prefix(rng, N, k)))
    return out


def poisson_requests(n: int, lam: float, seed: int = 0x5E7E, n_range=(12, 19), k_range=(0, 5)):
    """[(arrival_s, N, cols), ...] with exponential inter-arrival times of mean 1/lam."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / lam))
        N = int(rng.integers(n_range[0], n_range[1] + 1))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        out.append((t, N, random_prefix(rng, N, k)))
    return out


def instances_digest(insts) -> str:
    """A short digest of an instance list [(N, cols), ...]: tells whether a stored per-instance
    result (tests/golden/instance_counts.json) was made for exactly these instances."""
    text = ";".join(f"{N}:{','.join(map(str, cols))}" for N, cols in insts)
    return hashlib.sha256(text.encode()).hexdigest()[:16]
