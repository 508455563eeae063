"""Pin the CPU oracle before trusting it as the GPU checker.

Each oracle function is checked against something that does not restate its own formula:
the reference's published counts and log task counts, vectors produced by the reference's own
CPU build, an independent permutation brute force, and structural invariants.
"""
import itertools
import os

import numpy as np
import pytest

import oracle


def brute_solutions(N: int) -> int:
    """Permutation brute force: queens in distinct rows/columns, no shared diagonal."""
    n = 0
    for perm in itertools.permutations(range(N)):
        if len({i + c for i, c in enumerate(perm)}) == N and len({i - c for i, c in enumerate(perm)}) == N:
            n += 1
    return n


def brute_partials(N: int, depth: int) -> int:
    """Number of non-attacking placements of queens in rows 0..depth-1 (brute force)."""
    n = 0
    for cols in itertools.permutations(range(N), depth):
        if len({i + c for i, c in enumerate(cols)}) == depth and len({i - c for i, c in enumerate(cols)}) == depth:
            n += 1
    return n


def attacks(cols) -> bool:
    for i in range(len(cols)):
        for j in range(i):
            if cols[i] == cols[j] or abs(cols[i] - cols[j]) == i - j:
                return True
    return False


@pytest.mark.parametrize("N", range(1, 10))
def test_count_subtree_vs_brute_force(N):
    assert oracle.count_subtree(N, 0, 0, 0) == brute_solutions(N)


@pytest.mark.parametrize("N", range(4, 17))
def test_parallel_totals_vs_published(N, published):
    rows = min(5, N - 2)
    assert oracle.parallel_n_queens(N, rows, threads=8) == published["solutions"][N]


@pytest.mark.parametrize("N", range(4, 13))
def test_serial_every_rows(N, published):
    # 2 * sum over the symmetric task list equals the full count for every legal rows
    # (rows >= 2: the odd-N middle-column half needs its second row restricted).
    for rows in range(2, N - 1):
        assert oracle.serial_n_queens(N, rows) == published["solutions"][N], rows


def test_ref_vectors(ref_vectors):
    """Oracle task lists and per-task counts equal the reference's own CPU output."""
    for N, rows, total in ref_vectors["totals"]:
        N, rows, total = int(N), int(rows), int(total)
        t = oracle.gen_tasks(N, rows)
        assert np.array_equal(t, ref_vectors[f"N{N}_r{rows}_tasks"]), (N, rows)
        s, pt = oracle.count_tasks(N, t, per_task=True)
        assert np.array_equal(pt, ref_vectors[f"N{N}_r{rows}_counts"]), (N, rows)
        assert 2 * s == total


def test_task_counts_vs_logs(published):
    """Generator sizes vs the reference logs (rows 6, N=20..24; larger ones in the slow test)."""
    for N, rows, k, _src in published["task_counts"]:
        if N <= 24:
            assert oracle.lib().oracle_gen_tasks(N, rows, None, 0) == k, (N, rows)


@pytest.mark.slow
def test_task_counts_vs_logs_large(published):
    for N, rows, k, _src in published["task_counts"]:
        if 24 < N <= 26:
            assert oracle.lib().oracle_gen_tasks(N, rows, None, 0) == k, (N, rows)


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (no reference checkout)")
def test_gen_vs_reference_build():
    for N in range(4, 17):
        for rows in (2, 3, 4, 5):
            if rows <= N - 2:
                assert np.array_equal(oracle.gen_tasks(N, rows), oracle.ref_gen_tasks(N, rows)), (N, rows)


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (no reference checkout)")
def test_counts_vs_reference_build():
    R = oracle.ref()
    rng = np.random.default_rng(7)
    t = oracle.gen_tasks(14, 4)
    for i in rng.choice(len(t), 50, replace=False):
        c, l, r = (int(x) for x in t[i])
        assert oracle.count_subtree(14, c, l, r) == R.ref_count_subtree(14, c, l, r)
    assert oracle.parallel_n_queens(15, 5, 8) == R.ref_serial_n_queens(15, 5)


def test_generator_invariants():
    for N in (8, 9, 12, 13):
        for rows in (2, 3, 4):
            t = oracle.gen_tasks(N, rows)
            last = (1 << N) - 1
            for c, l, r in t:
                c, l, r = int(c), int(l), int(r)
                assert bin(c).count("1") == rows
                assert c & ~last == 0
        # Row 0 is a column < N/2, or (odd N) the middle column.
        cols0 = [int(c).bit_length() - 1 for c, _, _ in oracle.gen_tasks(N, 1)]
        assert cols0 == list(range(N // 2)) + ([N // 2] if N % 2 else [])
        # Every task is a non-attacking placement: the masks decode to `rows` queens.
        for c, l, r in oracle.gen_tasks(N, 3):
            assert bin(int(c)).count("1") == 3


@pytest.mark.parametrize("N", range(4, 9))
def test_count_nodes_vs_brute_force(N):
    # Nodes below the empty board = non-attacking partial placements at depths 1..N-1.
    expect = sum(brute_partials(N, d) for d in range(1, N))
    assert oracle.count_nodes(N, 0, 0, 0) == expect
    # Below a depth-2 prefix: placements at depths 3..N-1 that extend it.
    pref = oracle.encode_prefix(N, [0, 2])
    if pref is not None:
        n = 0
        for d in range(3, N):
            for rest in itertools.permutations([c for c in range(N) if c not in (0, 2)], d - 2):
                if not attacks([0, 2, *rest]):
                    n += 1
        assert oracle.count_nodes(N, *pref) == n


def test_encode_prefix_vs_attack_check():
    rng = np.random.default_rng(11)
    for _ in range(300):
        N = int(rng.integers(4, 14))
        k = int(rng.integers(0, N + 1))
        cols = [int(x) for x in rng.integers(0, N, size=k)]
        enc = oracle.encode_prefix(N, cols)
        assert (enc is None) == attacks(cols)
        if enc is not None and k < N:
            # completions of an encoded prefix == brute-force completions for small N
            if N <= 8:
                free = [c for c in range(N) if c not in cols]
                n = sum(1 for rest in itertools.permutations(free) if not attacks([*cols, *rest]))
                assert oracle.count_subtree(N, *enc) == n


def _random_prefixes(rng, n, N_range, k_max_gap=0):
    """Seeded non-attacking prefixes: N in N_range, k in 0..N (every depth, full boards too)."""
    out = []
    while len(out) < n:
        N = int(rng.integers(*N_range))
        k = int(rng.integers(0, N + 1 - k_max_gap))
        cols = []
        for row in range(k):
            free = [c for c in range(N) if not attacks([*cols, c])]
            if not free:
                break
            cols.append(int(free[rng.integers(len(free))]))
        if len(cols) == k:
            out.append((N, cols, oracle.encode_prefix(N, cols)))
    return out


def test_count_iterative_against_the_recursive_count_and_its_edge_cases():
    """Below N-1 queens the iterative variant counts exactly the completions (brute force for
    N <= 8, the recursion otherwise); a prefix of N-1 queens or a full board adds nothing,
    which is what the reference's loop does (src/n_queens.cpp:29-31,56-66)."""
    rng = np.random.default_rng(0x17E5)
    seen_edge = 0
    for N, cols, enc in _random_prefixes(rng, 400, (4, 13)):
        got = oracle.count_iterative(N, *enc)
        if len(cols) >= N - 1:
            assert got == 0, (N, cols)
            seen_edge += 1
        else:
            assert got == oracle.count_subtree(N, *enc), (N, cols)
            if N <= 8:
                free = [c for c in range(N) if c not in cols]
                assert got == sum(1 for rest in itertools.permutations(free) if not attacks([*cols, *rest]))
    assert seen_edge > 0
    # The whole board: A000170 (n_queens_iterative is the reference's "a little slower" counter).
    for N, want in ((8, 92), (10, 724), (12, 14200)):
        assert oracle.count_iterative(N, 0, 0, 0) == want


@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (no reference checkout)")
def test_count_iterative_vs_reference_build():
    """Against the reference's own n_queens_iterative, within its 16-frame stack (prefixes at
    most 17 rows above the last row; N <= 14 here)."""
    R = oracle.ref()
    rng = np.random.default_rng(0x1E7A)
    for N, cols, enc in _random_prefixes(rng, 300, (2, 15)):
        assert oracle.count_iterative(N, *enc) == R.ref_count_iterative(N, *enc), (N, cols)
