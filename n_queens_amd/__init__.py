"""MI355X-native N-Queens solution counter — Python host mirror of the C ABI (include/nq.h).

The product is ``libnq_hip.so`` (HIP kernels for gfx950 + C ABI + C++ drop-in) and the
``n_queens`` CLI, both built in-tree by ``n_queens_amd.build``.  This module only binds the C
ABI with ctypes; it never computes counts itself and raises when the library is missing.

Reference interface mirrored (ygch/n_queens, src/n_queens.h:7-19):

* ``cuda_n_queens(N, rows)``            -> ``long long cuda_n_queens(int N, int rows)``
* ``partial_n_queens(N, rows)``         -> task list of ``partial_n_queens`` (+ ``_for_odd``)
* ``count(N, tasks)``                   -> the per-GPU region of ``cuda_n_queens``
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnq_hip.so")
CLI_PATH = os.path.join(_HERE, "n_queens")

NQ_OK = 0
NQ_ERR_ARG = -1
NQ_ERR_NO_DEVICE = -2
NQ_ERR_HIP = -3
NQ_ERR_DEPTH = -4
NQ_ERR_NOMEM = -5
NQ_ERR_BUSY = -6
NQ_ERR_STATE = -7
NQ_PUSH_AUTO, NQ_PUSH_MASKED, NQ_PUSH_ALWAYS = 0, 1, 2

# Every symbol include/nq.h declares (checked by tests/test_host.py).
EXPORTS = (
    "nq_device_count", "nq_last_error", "nq_set_device", "nq_device_info", "nq_config_default",
    "nq_generate", "nq_generate_device", "nq_count", "nq_count_device", "nq_solve", "nq_solve_config", "nq_config_resolve",
    "nq_batch_plan_create", "nq_batch_plan_launch", "nq_batch_plan_wait", "nq_batch_plan_run",
    "nq_batch_plan_destroy", "nq_batch_count", "nq_instance_count",
    "nq_server_config_default", "nq_server_open", "nq_server_submit", "nq_server_poll", "nq_server_wait",
    "nq_server_stats", "nq_server_timing", "nq_server_close", "nq_solve_resumable", "nq_count_static",
)
NQ_STATIC_MAX_TASKS = 1 << 20

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class NQError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"nq status {status}: {msg}")
        self.status = status


class _Config(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("stack_depth", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("chunk", ctypes.c_int), ("grid", ctypes.c_int), ("split", ctypes.c_int), ("push", ctypes.c_int),
                ("preplace", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("ntasks", ctypes.c_uint64), ("nodes", ctypes.c_uint64), ("gen_ms", ctypes.c_double),
                ("h2d_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("busy_frac", ctypes.c_double), ("tail_ms", ctypes.c_double), ("waves", ctypes.c_uint64)]


class _Instance(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("k", ctypes.c_int), ("cols", ctypes.c_int8 * 32)]


class _BatchStats(ctypes.Structure):
    _fields_ = [("ntasks", ctypes.c_uint64), ("total", ctypes.c_uint64), ("expand_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int), ("slot_tasks", ctypes.c_uint64), ("slot_requests", ctypes.c_int),
                ("max_pending_tasks", ctypes.c_uint64), ("depth_target", ctypes.c_int), ("block", ctypes.c_int),
                ("chunk", ctypes.c_int), ("batch_work", ctypes.c_double)]


class _RequestStats(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("ntasks", ctypes.c_uint64)]


class _ResumeState(ctypes.Structure):
    _fields_ = [("ntasks", ctypes.c_uint64), ("next", ctypes.c_uint64), ("sum", ctypes.c_uint64),
                ("total", ctypes.c_uint64), ("done", ctypes.c_int), ("ranges", ctypes.c_int),
                ("gen_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double)]


@dataclass
class Config:
    """Launch configuration.  See nq_config in include/nq.h: 0 = library default for every
    field except split, where 0 means no device split and -1 (the default here) the library's
    automatic choice, as solve(N, rows) without a Config makes it."""
    block: int = 0
    stack_depth: int = 0
    waves_per_cu: int = 0
    chunk: int = 0
    grid: int = 0
    split: int = -1
    push: int = 0
    preplace: int = 0

    def _c(self) -> _Config:
        return _Config(self.block, self.stack_depth, self.waves_per_cu, self.chunk, self.grid, self.split,
                       self.push, self.preplace)

    @staticmethod
    def _from(c: _Config) -> "Config":
        return Config(c.block, c.stack_depth, c.waves_per_cu, c.chunk, c.grid, c.split, c.push, c.preplace)


_lib = None


def lib() -> ctypes.CDLL:
    """Load libnq_hip.so.  Raises if it has not been built (there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m n_queens_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.nq_device_count.restype = ctypes.c_int
    L.nq_last_error.restype = ctypes.c_char_p
    L.nq_set_device.argtypes = [ctypes.c_int]
    L.nq_device_info.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.c_char_p, ctypes.c_size_t]
    L.nq_config_default.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Config)]
    L.nq_solve_config.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Config)]
    L.nq_config_resolve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_Config),
                                    ctypes.POINTER(_Config)]
    L.nq_generate.argtypes = [ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_uint64, _u64p]
    L.nq_generate_device.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, _u64p,
                                     ctypes.c_void_p]
    L.nq_count.argtypes = [ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.POINTER(_Config), _u64p, _u64p, _u64p,
                           ctypes.POINTER(_Stats)]
    L.nq_count_device.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Config),
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.nq_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Config), _u64p, ctypes.POINTER(_Stats)]
    L.nq_batch_plan_create.argtypes = [ctypes.POINTER(_Instance), ctypes.c_int, ctypes.POINTER(_Config),
                                       ctypes.POINTER(ctypes.c_void_p)]
    L.nq_batch_plan_launch.argtypes = [ctypes.c_void_p]
    L.nq_batch_plan_wait.argtypes = [ctypes.c_void_p, _u64p, ctypes.POINTER(_BatchStats)]
    L.nq_batch_plan_run.argtypes = [ctypes.c_void_p, _u64p, ctypes.POINTER(_BatchStats)]
    L.nq_batch_plan_destroy.argtypes = [ctypes.c_void_p]
    L.nq_batch_count.argtypes = [ctypes.POINTER(_Instance), ctypes.c_int, ctypes.POINTER(_Config), _u64p,
                                 ctypes.POINTER(_BatchStats)]
    L.nq_instance_count.argtypes = [ctypes.POINTER(_Instance), ctypes.POINTER(_Config), _u64p]
    L.nq_server_config_default.argtypes = [ctypes.POINTER(_ServerConfig)]
    L.nq_server_open.argtypes = [ctypes.POINTER(_ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
    L.nq_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Instance), _u64p]
    L.nq_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, _u64p, ctypes.POINTER(_RequestStats)]
    L.nq_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, _u64p, ctypes.POINTER(_RequestStats)]
    L.nq_server_stats.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p]
    L.nq_server_close.argtypes = [ctypes.c_void_p]
    L.nq_server_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.nq_solve_resumable.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Config), ctypes.c_char_p,
                                     ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_ResumeState)]
    L.nq_count_static.argtypes = [ctypes.c_int, _u32p, ctypes.c_uint64, ctypes.POINTER(_Config), _u64p, _u64p,
                                  _u64p]
    for name in EXPORTS:
        if name not in ("nq_last_error",):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != NQ_OK:
        raise NQError(rc, lib().nq_last_error().decode(errors="replace"))


def device_count() -> int:
    return int(lib().nq_device_count())


def device_info(device: int = 0) -> dict:
    cus, clk = ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(256)
    _check(lib().nq_device_info(device, ctypes.byref(cus), ctypes.byref(clk), name, 256))
    return {"cus": cus.value, "clock_khz": clk.value, "name": name.value.decode()}


def config_default(N: int, d0: int) -> Config:
    c = _Config()
    _check(lib().nq_config_default(N, d0, ctypes.byref(c)))
    return Config._from(c)


def solve_config(N: int, rows: int) -> Config:
    """The configuration nq_solve uses for (N, rows), including its device split."""
    c = _Config()
    _check(lib().nq_solve_config(N, rows, ctypes.byref(c)))
    return Config._from(c)


def resolve_config(N: int, d0: int, ntasks: int, cfg: Config | None = None) -> Config:
    """The launch configuration nq_count takes on this device for `ntasks` tasks of shallowest
    prefix depth d0 (after any split): workgroup size, push form and grid resolved."""
    c = _Config()
    cc = cfg._c() if cfg else None
    _check(lib().nq_config_resolve(N, d0, ntasks, ctypes.byref(cc) if cc else None, ctypes.byref(c)))
    return Config._from(c)


def generate(N: int, rows: int) -> np.ndarray:
    """Task list (ntasks, 3) uint32 of partial_n_queens + partial_n_queens_for_odd."""
    L = lib()
    n = ctypes.c_uint64()
    _check(L.nq_generate(N, rows, None, 0, ctypes.byref(n)))
    out = np.zeros((max(n.value, 1), 3), dtype=np.uint32)
    _check(L.nq_generate(N, rows, out.ctypes.data_as(_u32p), n.value, ctypes.byref(n)))
    return out[: n.value]


partial_n_queens = generate


def generate_device(N: int, rows: int, d_tasks: int = 0, cap: int = 0, stream: int = 0) -> int:
    """nq_generate_device: the same task list generated on the current device into the device
    buffer at d_tasks (an int pointer to 3 * cap uint32), synchronously on `stream`.  Returns
    the task count; with d_tasks 0 or cap too small nothing is written."""
    n = ctypes.c_uint64()
    _check(lib().nq_generate_device(N, rows, ctypes.c_void_p(d_tasks or None), cap, ctypes.byref(n),
                                    ctypes.c_void_p(stream or None)))
    return n.value


@dataclass
class CountResult:
    sum: int
    per_task: np.ndarray | None
    nodes: int | None
    kernel_ms: float
    h2d_ms: float
    total_ms: float
    busy_frac: float = 0.0
    tail_ms: float = 0.0
    waves: int = 0


def count(N: int, tasks: np.ndarray, cfg: Config | None = None, per_task: bool = False,
          nodes: bool = False) -> CountResult:
    """Count completions of every task on the GPU (nq_count)."""
    t = np.ascontiguousarray(tasks, dtype=np.uint32).reshape(-1, 3)
    n = t.shape[0]
    s = ctypes.c_uint64()
    nd = ctypes.c_uint64()
    pt = np.zeros(max(n, 1), dtype=np.uint64) if per_task else None
    st = _Stats()
    cc = cfg._c() if cfg else None
    _check(lib().nq_count(N, t.ctypes.data_as(_u32p), n, ctypes.byref(cc) if cc else None, ctypes.byref(s),
                          pt.ctypes.data_as(_u64p) if per_task else None, ctypes.byref(nd) if nodes else None,
                          ctypes.byref(st)))
    return CountResult(int(s.value), pt[:n] if per_task else None, int(nd.value) if nodes else None,
                       st.kernel_ms, st.h2d_ms, st.total_ms, st.busy_frac, st.tail_ms, int(st.waves))


def count_static(N: int, tasks: np.ndarray, cfg: Config | None = None, per_task: bool = False,
                 nodes: bool = False) -> CountResult:
    """nq_count_static: count() with no device allocation (at most NQ_STATIC_MAX_TASKS tasks)."""
    t = np.ascontiguousarray(tasks, dtype=np.uint32).reshape(-1, 3)
    n = t.shape[0]
    s = ctypes.c_uint64()
    nd = ctypes.c_uint64()
    pt = np.zeros(max(n, 1), dtype=np.uint64) if per_task else None
    cc = cfg._c() if cfg else None
    _check(lib().nq_count_static(N, t.ctypes.data_as(_u32p), n, ctypes.byref(cc) if cc else None, ctypes.byref(s),
                                 pt.ctypes.data_as(_u64p) if per_task else None,
                                 ctypes.byref(nd) if nodes else None))
    return CountResult(int(s.value), pt[:n] if per_task else None, int(nd.value) if nodes else None, 0.0, 0.0, 0.0)


def count_device(N: int, d_tasks: int, ntasks: int, d_result: int, d_per_task: int = 0, cfg: Config | None = None,
                 count_nodes: bool = False, stream: int = 0) -> None:
    """Enqueue nq_count_device on raw device pointers (ints) and a hipStream_t (int)."""
    cc = cfg._c() if cfg else None
    _check(lib().nq_count_device(N, ctypes.c_void_p(d_tasks), ntasks, ctypes.byref(cc) if cc else None,
                                 ctypes.c_void_p(d_result), ctypes.c_void_p(d_per_task or None),
                                 1 if count_nodes else 0, ctypes.c_void_p(stream or None)))


def solve(N: int, rows: int, cfg: Config | None = None) -> tuple[int, dict]:
    """Number of N-Queens solutions with `rows` pre-placed (nq_solve = cuda_n_queens)."""
    tot = ctypes.c_uint64()
    st = _Stats()
    cc = cfg._c() if cfg else None
    _check(lib().nq_solve(N, rows, ctypes.byref(cc) if cc else None, ctypes.byref(tot), ctypes.byref(st)))
    return int(tot.value), {"ntasks": st.ntasks, "gen_ms": st.gen_ms, "h2d_ms": st.h2d_ms,
                            "kernel_ms": st.kernel_ms, "total_ms": st.total_ms, "busy_frac": st.busy_frac,
                            "tail_ms": st.tail_ms, "waves": int(st.waves)}


def cuda_n_queens(N: int, rows: int) -> int:
    """Reference-compatible return value: the solution count, or -1 on failure."""
    try:
        return solve(N, rows)[0]
    except NQError:
        return -1


# ---- batched instances -------------------------------------------------------------------

def _instances(insts) -> ctypes.Array:
    """[(N, [cols...]), ...] -> nq_instance array."""
    arr = (_Instance * max(1, len(insts)))()
    for i, (N, cols) in enumerate(insts):
        arr[i].N = int(N)
        arr[i].k = len(cols)
        for j, c in enumerate(cols):
            arr[i].cols[j] = int(c)
    return arr


def batch_count(insts, cfg: Config | None = None) -> tuple[np.ndarray, dict]:
    """Counts of many (N, prefix-columns) instances in one batched launch (nq_batch_count)."""
    arr = _instances(insts)
    out = np.zeros(max(1, len(insts)), dtype=np.uint64)
    st = _BatchStats()
    cc = cfg._c() if cfg else None
    _check(lib().nq_batch_count(arr, len(insts), ctypes.byref(cc) if cc else None, out.ctypes.data_as(_u64p),
                                ctypes.byref(st)))
    return out[: len(insts)], {"ntasks": st.ntasks, "total": st.total, "expand_ms": st.expand_ms,
                               "kernel_ms": st.kernel_ms, "total_ms": st.total_ms}


def solve_resumable(N: int, rows: int, path: str, range_tasks: int = 0, max_ranges: int = 0,
                    cfg: Config | None = None) -> dict:
    """nq_solve_resumable: count (N, rows) in task ranges, checkpointing to `path` after each
    range and resuming from it.  Returns the state: total (2 x sum) once done."""
    st = _ResumeState()
    cc = cfg._c() if cfg else None
    _check(lib().nq_solve_resumable(N, rows, ctypes.byref(cc) if cc else None, os.fsencode(path), range_tasks,
                                    max_ranges, ctypes.byref(st)))
    return {"ntasks": st.ntasks, "next": st.next, "sum": st.sum, "total": st.total, "done": bool(st.done),
            "ranges": st.ranges, "gen_ms": st.gen_ms, "kernel_ms": st.kernel_ms}


def instance_count(N: int, cols, cfg: Config | None = None) -> int:
    """One instance through the plain per-instance launch (nq_instance_count)."""
    arr = _instances([(N, cols)])
    out = ctypes.c_uint64()
    cc = cfg._c() if cfg else None
    _check(lib().nq_instance_count(arr, ctypes.byref(cc) if cc else None, ctypes.byref(out)))
    return int(out.value)


class BatchPlan:
    """Resident batch: expanded + uploaded once, launch sequence replayed from a hipGraph."""

    def __init__(self, insts, cfg: Config | None = None):
        self.n = len(insts)
        arr = _instances(insts)
        h = ctypes.c_void_p()
        cc = cfg._c() if cfg else None
        _check(lib().nq_batch_plan_create(arr, self.n, ctypes.byref(cc) if cc else None, ctypes.byref(h)))
        self._h = h

    def launch(self) -> None:
        _check(lib().nq_batch_plan_launch(self._h))

    def wait(self) -> tuple[np.ndarray, dict]:
        out = np.zeros(max(1, self.n), dtype=np.uint64)
        st = _BatchStats()
        _check(lib().nq_batch_plan_wait(self._h, out.ctypes.data_as(_u64p), ctypes.byref(st)))
        return out[: self.n], {"ntasks": st.ntasks, "total": st.total, "expand_ms": st.expand_ms,
                               "kernel_ms": st.kernel_ms}

    def run(self) -> tuple[np.ndarray, dict]:
        self.launch()
        return self.wait()

    def close(self) -> None:
        if self._h:
            lib().nq_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- serving -------------------------------------------------------------------------------

class Server:
    """Continuous-batching request server (nq_server_*)."""

    def __init__(self, **cfg):
        c = _ServerConfig()
        _check(lib().nq_server_config_default(ctypes.byref(c)))
        for k, v in cfg.items():
            setattr(c, k, v)
        h = ctypes.c_void_p()
        _check(lib().nq_server_open(ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def submit(self, N: int, cols) -> int:
        arr = _instances([(N, cols)])
        t = ctypes.c_uint64()
        _check(lib().nq_server_submit(self._h, arr, ctypes.byref(t)))
        return int(t.value)

    def poll(self, ticket: int):
        """(count, stats) when done, None while pending."""
        c = ctypes.c_uint64()
        st = _RequestStats()
        rc = lib().nq_server_poll(self._h, ticket, ctypes.byref(c), ctypes.byref(st))
        if rc == NQ_ERR_STATE:
            return None
        _check(rc)
        return int(c.value), {"queue_ms": st.queue_ms, "latency_ms": st.latency_ms, "ntasks": int(st.ntasks)}

    def wait(self, ticket: int, timeout_ms: int = -1):
        c = ctypes.c_uint64()
        st = _RequestStats()
        _check(lib().nq_server_wait(self._h, ticket, timeout_ms, ctypes.byref(c), ctypes.byref(st)))
        return int(c.value), {"queue_ms": st.queue_ms, "latency_ms": st.latency_ms, "ntasks": int(st.ntasks)}

    def stats(self) -> dict:
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().nq_server_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        d, o = ctypes.c_double(), ctypes.c_double()
        _check(lib().nq_server_timing(self._h, ctypes.byref(d), ctypes.byref(o)))
        return {"completed": a.value, "batches": b.value, "solutions": c.value, "device_ms": d.value,
                "open_ms": o.value}

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            _check(lib().nq_server_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
