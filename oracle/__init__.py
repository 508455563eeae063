"""CPU oracle for the N-Queens counting path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker or the timed CPU baseline.  The product
(``n_queens_amd``) never imports it.

``lib()`` loads ``oracle/libnq_oracle.so`` (this repo's C restatement of the reference's CPU
algorithms, ``oracle/nq_oracle.c``).  ``ref()`` loads ``oracle/_ref/libnq_ref.so`` when it was
built — the reference's own ``src/n_queens.cpp`` compiled where it lies by ``oracle/Makefile``
— and returns None otherwise.

Parity status of each oracle function (how it is pinned, see ``tests/test_oracle.py``):

* ``gen_tasks``      — task counts vs the reference logs (N=20..26 rows 6, N=27 rows 7),
                       exact task lists vs ``_ref`` (N=4..16), golden fixtures, invariants.
* ``count_subtree``  — totals vs OEIS A000170 (README.md:16-25, logs), permutation brute force
                       N<=9, per-task counts vs ``_ref`` and golden fixtures.
* ``count_iterative``— the reference's iterative variant: equal to ``_ref`` on random
                       prefixes, to ``count_subtree`` below N-1 queens, 0 at N-1 and N queens.
* ``count_nodes``    — brute-force recursion with an independent definition (placements at
                       depths rows+1..N-1), closed form for the full tree on tiny N.
* ``serial/parallel``— A000170 and ``_ref``.
* ``encode_prefix``  — brute-force attack check on random prefixes.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libnq_oracle.so")
_REF = os.path.join(_HERE, "_ref", "libnq_ref.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None
_ref = None


def build(ref_src: str = "/root/reference/src") -> None:
    """Compile the oracle (and the reference CPU sources into _ref/ when they exist)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "all"])
    if os.path.isdir(ref_src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "ref", f"REF_SRC={ref_src}"])


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_gen_tasks.restype = ctypes.c_uint64
        L.oracle_gen_tasks.argtypes = [ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_uint64]
        for name in ("oracle_count_subtree", "oracle_count_nodes", "oracle_count_iterative"):
            f = getattr(L, name)
            f.restype = ctypes.c_uint64
            f.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_count_tasks.restype = ctypes.c_uint64
        L.oracle_count_tasks.argtypes = [ctypes.c_int, _u32p, ctypes.c_uint64, _u64p, ctypes.c_int]
        L.oracle_count_tasks_nodes.restype = ctypes.c_uint64
        L.oracle_count_tasks_nodes.argtypes = [ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.c_int]
        L.oracle_serial_n_queens.restype = ctypes.c_uint64
        L.oracle_serial_n_queens.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_parallel_n_queens.restype = ctypes.c_uint64
        L.oracle_parallel_n_queens.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_encode_prefix.restype = ctypes.c_int
        L.oracle_encode_prefix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           _u32p, _u32p, _u32p]
        _lib = L
    return _lib


def ref():
    """The reference's own CPU functions, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        R = ctypes.CDLL(_REF)
        R.ref_gen_tasks.restype = ctypes.c_uint64
        R.ref_gen_tasks.argtypes = [ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_uint64]
        R.ref_count_subtree.restype = ctypes.c_uint64
        R.ref_count_subtree.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        R.ref_count_iterative.restype = ctypes.c_uint64
        R.ref_count_iterative.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        R.ref_serial_n_queens.restype = ctypes.c_uint64
        R.ref_serial_n_queens.argtypes = [ctypes.c_int, ctypes.c_int]
        R.ref_parallel_n_queens.restype = ctypes.c_uint64
        R.ref_parallel_n_queens.argtypes = [ctypes.c_int, ctypes.c_int]
        _ref = R
    return _ref


def _p32(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def gen_tasks(N: int, rows: int) -> np.ndarray:
    """Task list as a uint32 array of shape (ntasks, 3): (cur, left, right)."""
    L = lib()
    n = L.oracle_gen_tasks(N, rows, None, 0)
    out = np.zeros((max(n, 1), 3), dtype=np.uint32)
    L.oracle_gen_tasks(N, rows, _p32(out), n)
    return out[:n]


def ref_gen_tasks(N: int, rows: int) -> np.ndarray:
    R = ref()
    n = R.ref_gen_tasks(N, rows, None, 0)
    out = np.zeros((max(n, 1), 3), dtype=np.uint32)
    R.ref_gen_tasks(N, rows, _p32(out), n)
    return out[:n]


def count_subtree(N: int, cur: int, left: int, right: int) -> int:
    return int(lib().oracle_count_subtree(N, cur & 0xFFFFFFFF, left & 0xFFFFFFFF, right & 0xFFFFFFFF))


def count_iterative(N: int, cur: int, left: int, right: int) -> int:
    return int(lib().oracle_count_iterative(N, cur & 0xFFFFFFFF, left & 0xFFFFFFFF, right & 0xFFFFFFFF))


def count_nodes(N: int, cur: int, left: int, right: int) -> int:
    return int(lib().oracle_count_nodes(N, cur & 0xFFFFFFFF, left & 0xFFFFFFFF, right & 0xFFFFFFFF))


def count_tasks(N: int, tasks: np.ndarray, threads: int = 0, per_task: bool = False):
    """Sum of completions over a (n, 3) task array; with per_task also the uint64 vector."""
    t = np.ascontiguousarray(tasks, dtype=np.uint32)
    n = t.shape[0]
    pt = np.zeros(max(n, 1), dtype=np.uint64) if per_task else None
    s = lib().oracle_count_tasks(N, _p32(t), n, pt.ctypes.data_as(_u64p) if per_task else None, threads)
    return (int(s), pt[:n]) if per_task else int(s)


def count_tasks_nodes(N: int, tasks: np.ndarray, threads: int = 0) -> int:
    t = np.ascontiguousarray(tasks, dtype=np.uint32)
    return int(lib().oracle_count_tasks_nodes(N, _p32(t), t.shape[0], threads))


def serial_n_queens(N: int, rows: int) -> int:
    return int(lib().oracle_serial_n_queens(N, rows))


def parallel_n_queens(N: int, rows: int, threads: int = 0) -> int:
    return int(lib().oracle_parallel_n_queens(N, rows, threads))


def encode_prefix(N: int, cols) -> tuple[int, int, int] | None:
    k = len(cols)
    arr = (ctypes.c_int * max(k, 1))(*cols)
    c, l, r = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    rc = lib().oracle_encode_prefix(N, k, arr, ctypes.byref(c), ctypes.byref(l), ctypes.byref(r))
    return None if rc != 0 else (c.value, l.value, r.value)
