"""The DFS hot loop's own instruction text, run on the CPU — test infrastructure only.

tests/loop_model.c restates the loop in C; this module runs the loop as the product ships it.
It preprocesses n_queens_amd/csrc/nq_dfs.hip for gfx950 with the HIP compiler (so every macro is
expanded exactly as the device build expands it), takes the six `__asm__` blocks of `hot_loop`
(masked or unconditional push x per-wave count, per-lane count, per-lane count with node
counting), and interprets their instructions for one wave of 64 lanes with numpy: VGPRs, the
exec mask, vcc, SGPRs, SCC and the workgroup's LDS.  A driver around it does what dfs_kernel does
around hot_loop (nq_dfs.hip): lanes parked on the dead frame claim the next task, the frame
registers and stack pointer are set as the kernel sets them, the loop runs for `live`, and the
32-bit counters are folded after each return.

Checks made while interpreting:
  * every ds_write_b128 / ds_read_b128 of a lane touches that lane's own 16-byte slot at a level
    below `levels`, or the dead frame at address 0 (and a write there leaves it unchanged);
  * the loop returns within a bound of trips, every lane that ends a task is parked on the dead
    frame with no candidate, and no instruction outside the known set appears.

Only the instructions the loop uses are modelled; anything else raises.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "n_queens_amd", "csrc", "nq_dfs.hip")
M32 = np.uint64(0xFFFFFFFF)
LANES = 64
_SHIFTS = np.arange(64, dtype=np.uint64)


def _popcount(x: np.ndarray) -> np.ndarray:
    x = x & M32
    x = x - ((x >> np.uint64(1)) & np.uint64(0x55555555))
    x = (x & np.uint64(0x33333333)) + ((x >> np.uint64(2)) & np.uint64(0x33333333))
    x = (x + (x >> np.uint64(4))) & np.uint64(0x0F0F0F0F)
    return ((x * np.uint64(0x01010101)) & M32) >> np.uint64(24)
DEAD = (0xFFFFFFFE, 0, 0xFFFFFFFF, 0)  # put_dead_frame() in nq_dfs.hip


def compiler() -> str | None:
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    return None


def preprocess() -> str:
    """nq_dfs.hip preprocessed for the device (gfx950), as text."""
    cxx = compiler()
    if cxx is None:
        raise RuntimeError("no HIP compiler")
    cmd = [cxx, "-x", "hip", "--cuda-device-only", "--offload-arch=gfx950", "-nogpulib", "-E", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "n_queens_amd", "csrc"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", SRC]
    return subprocess.run(cmd, check=True, capture_output=True, text=True).stdout


def _strings(s: str) -> str:
    """Concatenation of the C string literals in s (the escapes the asm uses: \\n, \\t)."""
    out = []
    for m in re.finditer(r'"((?:[^"\\]|\\.)*)"', s):
        out.append(m.group(1).replace("\\n", "\n").replace("\\t", "\t"))
    return "".join(out)


def hot_loop_blocks(text: str) -> dict[str, list[str]]:
    """The asm blocks of hot_loop, keyed by variant: '<push>_<count>[_nodes]' with push in
    {always, masked} and count in {wave, lane}.  Each value is the instruction list (labels
    included, `%=` kept)."""
    i = text.index("void hot_loop(")
    j = text.index("struct LdsLayout", i)
    body = text[i:j]
    blocks = {}
    for m in re.finditer(r"__asm__ volatile\(", body):
        # The template: the string literals up to the first ':' outside a literal.
        k = m.end()
        lits = []
        while True:
            while body[k].isspace():
                k += 1
            if body[k] != '"':
                break
            e = k + 1
            while body[e] != '"':
                e += 2 if body[e] == "\\" else 1
            lits.append(body[k:e + 1])
            k = e + 1
        asm = _strings(" ".join(lits))
        lines = [ln.strip() for ln in asm.split("\n") if ln.strip()]
        push = "always" if any(ln.startswith("v_mad_u32_u24 %[sp]") for ln in lines) else "masked"
        count = "wave" if any(ln.startswith("s_bcnt1_i32_b64 %[bc], vcc") for ln in lines) else "lane"
        nodes = any("%[nd]" in ln.split(",")[0] for ln in lines)
        key = f"{push}_{count}" + ("_nodes" if nodes else "")
        blocks[key] = lines
    return blocks


class LoopError(AssertionError):
    pass


class Wave:
    """One wave's registers and its workgroup's LDS (uint32 words)."""

    def __init__(self, lds_words: int, levels: int, stride: int):
        self.v = {r: np.zeros(LANES, np.uint64) for r in ("v20", "v21", "v22", "v23")}
        self.named = {}           # VGPR operands by name: sp, cnt, nd, t, p
        self.s = {}               # SGPR operands by name: iters, wcnt, m1, save, bc, live, stride
        self.exec = 0
        self.vcc = 0
        self.scc = 0
        self.lds = np.zeros(lds_words, np.uint64)
        self.levels = levels
        self.stride = stride
        self.lane0 = 0            # the wave's first lane in its workgroup

    # -- operands -------------------------------------------------------------------------
    def vget(self, tok: str):
        tok = tok.strip()
        if tok.startswith("%["):
            name = tok[2:-1]
            if name in self.named:
                return self.named[name]
            return np.full(LANES, self.s[name] & 0xFFFFFFFF, np.uint64)  # an SGPR read as a VGPR operand
        if tok in self.v:
            return self.v[tok]
        return np.full(LANES, int(tok, 0) & 0xFFFFFFFF, np.uint64)

    def vset(self, tok: str, val, mask):
        tok = tok.strip()
        val = np.asarray(val, np.uint64) & M32
        if tok.startswith("%["):
            arr = self.named[tok[2:-1]]
        else:
            arr = self.v[tok]
        arr[mask] = val[mask]

    def sget(self, tok: str) -> int:
        tok = tok.strip()
        if tok == "exec":
            return self.exec
        if tok == "vcc":
            return self.vcc
        if tok.startswith("%["):
            return self.s[tok[2:-1]]
        return int(tok, 0) & 0xFFFFFFFFFFFFFFFF

    def sset(self, tok: str, val: int):
        tok = tok.strip()
        if tok == "exec":
            self.exec = val
        elif tok == "vcc":
            self.vcc = val
        else:
            self.s[tok[2:-1]] = val

    def active(self) -> np.ndarray:
        return ((np.uint64(self.exec) >> _SHIFTS) & np.uint64(1)).astype(bool)

    @staticmethod
    def to_mask(bits: np.ndarray) -> int:
        return int((np.asarray(bits, bool).astype(np.uint64) << _SHIFTS).sum())

    # -- LDS ------------------------------------------------------------------------------
    def owned(self, addr: int, lane: int) -> bool:
        if addr == 0:
            return True
        if addr < 16:
            return False
        off = addr - 16
        return off % self.stride == 16 * lane and off // self.stride < self.levels

    def lds_access(self, write: bool, addr_tok: str, act: np.ndarray):
        addr = self.vget(addr_tok)
        for i in np.nonzero(act)[0]:
            a = int(addr[i])
            if not self.owned(a, self.lane0 + int(i)) or a % 16:
                raise LoopError(f"{'write' if write else 'read'} of lane {i} at {a}: not its slot")
            w = a // 4
            if write:
                frame = tuple(int(self.v[r][i]) for r in ("v20", "v21", "v22", "v23"))
                if a == 0 and frame != DEAD:
                    raise LoopError(f"lane {i} wrote {frame} over the dead frame")
                self.lds[w:w + 4] = frame
            else:
                for k, r in enumerate(("v20", "v21", "v22", "v23")):
                    self.v[r][i] = self.lds[w + k]


def _ops(line: str) -> tuple[str, list[str], str]:
    op, _, rest = line.partition(" ")
    mod = ""
    for m in (" clamp", " bitop3:"):
        if m in rest:
            idx = rest.index(m)
            mod = rest[idx:].strip()
            rest = rest[:idx]
    return op, [a.strip() for a in rest.split(",") if a.strip()], mod


def run_block(w: Wave, lines: list[str], max_steps: int) -> None:
    """Interpret one asm block from its first instruction to its end."""
    labels = {ln[:-1]: k for k, ln in enumerate(lines) if ln.endswith(":")}
    pc = 0
    steps = 0
    while pc < len(lines):
        steps += 1
        if steps > max_steps:
            raise LoopError("the loop did not return within its step bound")
        ln = lines[pc]
        pc += 1
        if ln.endswith(":"):
            continue
        op, a, mod = _ops(ln)
        act = w.active()
        if op == "s_mov_b64":
            w.sset(a[0], w.sget(a[1]))
        elif op == "s_waitcnt":
            pass
        elif op == "s_bcnt1_i32_b64":
            n = bin(w.sget(a[1])).count("1")
            w.sset(a[0], n)
            w.scc = int(n != 0)
        elif op in ("s_add_u32", "s_sub_u32"):
            x, y = w.sget(a[1]) & 0xFFFFFFFF, w.sget(a[2]) & 0xFFFFFFFF
            r = x + y if op == "s_add_u32" else x - y
            w.scc = int(r > 0xFFFFFFFF or r < 0)
            w.sset(a[0], r & 0xFFFFFFFF)
        elif op in ("s_cmp_lg_u64", "s_cmp_lg_u32"):
            x, y = w.sget(a[0]), w.sget(a[1])
            if op.endswith("u32"):
                x, y = x & 0xFFFFFFFF, y & 0xFFFFFFFF
            w.scc = int(x != y)
        elif op == "s_cbranch_scc1":
            if w.scc:
                pc = labels[a[0]]
        elif op in ("ds_write_b128", "ds_read_b128"):
            if a[1 if op == "ds_write_b128" else 0] != "v[20:23]":
                raise LoopError(f"unexpected frame registers in {ln}")
            w.lds_access(op == "ds_write_b128", a[0] if op == "ds_write_b128" else a[1], act)
        elif op.startswith("v_cmp"):
            cmp = op.split("_")[2]  # eq / ne
            x, y = w.vget(a[1]), w.vget(a[2])
            res = (x == y) if cmp == "eq" else (x != y)
            mask = Wave.to_mask(res & act)
            w.sset(a[0], mask)
            if op.startswith("v_cmpx"):
                w.exec = mask
        elif op == "v_addc_co_u32":
            # D, carry-out, S0, S1, carry-in (vcc): inactive lanes keep D; carry-out 0 there.
            cin = (np.uint64(w.sget(a[4])) >> _SHIFTS) & np.uint64(1)
            r = w.vget(a[2]) + w.vget(a[3]) + cin
            w.vset(a[0], r, act)
            w.sset(a[1], Wave.to_mask((r > M32) & act))
        else:
            x = [w.vget(t) for t in a[1:]]
            if op == "v_sub_u32" or op == "v_sub_u32_e64":
                if mod == "clamp":
                    r = np.where(x[0] >= x[1], x[0] - x[1], 0)
                else:
                    r = (x[0] - x[1]) & M32
            elif op == "v_add_u32":
                r = x[0] + x[1]
            elif op == "v_and_b32":
                r = x[0] & x[1]
            elif op == "v_or_b32":
                r = x[0] | x[1]
            elif op == "v_xor_b32":
                r = x[0] ^ x[1]
            elif op == "v_min_u32":
                r = np.minimum(x[0], x[1])
            elif op == "v_mad_u32_u24":
                r = (x[0] & np.uint64(0xFFFFFF)) * (x[1] & np.uint64(0xFFFFFF)) + x[2]
            elif op == "v_lshrrev_b32":
                r = x[1] >> (x[0] & np.uint64(31))
            elif op == "v_add_lshl_u32":
                r = ((x[0] + x[1]) & M32) << (x[2] & np.uint64(31))
            elif op == "v_bcnt_u32_b32":
                r = _popcount(x[0]) + x[1]
            elif op == "v_bitop3_b32":
                tt = int(mod.split(":")[1], 0)
                r = np.zeros(LANES, np.uint64)
                for idx in range(8):
                    if (tt >> idx) & 1:
                        term = np.full(LANES, M32, np.uint64)
                        for k, src in enumerate(x):
                            bit = (idx >> (2 - k)) & 1
                            term &= src if bit else (~src & M32)
                        r |= term
            else:
                raise LoopError(f"instruction not modelled: {ln}")
            w.vset(a[0], r, act)


def count_tasks(lines: list[str], N: int, tasks: np.ndarray, push_always: bool, lane_count: bool, nodes: bool,
                frames: int, max_trips: int = 3, waves: int = 1, block: int = 64) -> tuple[np.ndarray, int, int]:
    """Count every task (3 words each) with the given asm block, as dfs_kernel runs hot_loop
    (lane-granular refill from a shared cursor, one frame per level per lane, a dead frame at
    address 0).  Returns (per-task counts, total, nodes); per-task counts only with a per-lane
    count (else zeros)."""
    last = (1 << N) - 1
    levels = frames + (1 if push_always else 0)
    stride = 16 * block
    lds_words = (16 + levels * stride) // 4
    ntasks = len(tasks) // 3 if tasks.ndim == 1 else len(tasks)
    tasks = tasks.reshape(-1, 3)
    per_task = np.zeros(ntasks, np.uint64)
    total = 0
    nodes_total = 0
    cursor = 0
    wv = []
    for k in range(waves):
        w = Wave(lds_words, levels, stride)
        w.lds[0:4] = DEAD
        w.named = {n: np.zeros(LANES, np.uint64) for n in ("sp", "cnt", "nd", "t", "p")}
        w.s = {"iters": 0, "wcnt": 0, "m1": 0, "save": 0, "bc": 0, "live": 0, "stride": stride}
        w.lane0 = 64 * k                  # this wave's first lane in the workgroup
        w.tid = np.full(LANES, -1, np.int64)
        w.has = np.ones(LANES, bool)
        w.tsum = np.zeros(LANES, np.uint64)
        wv.append(w)
        # wave k's lanes share the workgroup's LDS: give every wave the same array
        if k:
            w.lds = wv[0].lds
    for trip_round in range(10 ** 7):
        any_live = False
        for w in wv:
            v23 = w.v["v23"]
            for i in range(LANES):
                if w.has[i] and v23[i] == 0:
                    if w.tid[i] >= 0:
                        per_task[w.tid[i]] = w.tsum[i]
                        if int(w.named["sp"][i]) != 0:
                            raise LoopError(f"lane {i} ended a task off the dead frame")
                    total += int(w.tsum[i])
                    w.tsum[i] = 0
                    w.tid[i] = -1
                    w.has[i] = False
                    while cursor < ntasks:
                        t = cursor
                        cursor += 1
                        c0, l0, r0 = (int(x) for x in tasks[t])
                        v0 = last & ~(c0 | l0 | r0)
                        d0 = bin(c0).count("1")
                        if d0 + 1 >= N or v0 == 0:
                            direct = (1 if c0 == last else 0) if d0 >= N else (bin(v0).count("1") if d0 + 1 == N else 0)
                            per_task[t] = direct
                            total += direct
                            continue
                        if N - 1 - d0 > frames:
                            raise LoopError("task needs more frames than configured")
                        w.v["v20"][i] = (c0 | ~last) & 0xFFFFFFFF
                        w.v["v21"][i] = l0
                        w.v["v22"][i] = r0
                        w.v["v23"][i] = v0
                        w.named["sp"][i] = 16 + 16 * (w.lane0 + i)
                        w.tid[i] = t
                        w.has[i] = True
                        break
            live = Wave.to_mask(w.has)
            if not live:
                continue
            any_live = True
            w.exec = (1 << 64) - 1
            w.s["live"] = live
            w.s["iters"] = max_trips
            w.s["wcnt"] = 0
            w.named["cnt"][:] = 0
            w.named["nd"][:] = 0
            run_block(w, lines, max_steps=(max_trips + 1) * 40 * 16)
            if w.exec != (1 << 64) - 1:
                raise LoopError("exec not restored at the end of the loop")
            if lane_count:
                w.tsum += w.named["cnt"]
            else:
                total += w.s["wcnt"]
            if nodes:
                nodes_total += int((w.named["nd"] - w.named["cnt"]).sum())
        if not any_live:
            break
    else:
        raise LoopError("no termination")
    for w in wv:
        if (w.lds[0:4] != DEAD).any():
            raise LoopError("the dead frame changed")
    return per_task, total, nodes_total
