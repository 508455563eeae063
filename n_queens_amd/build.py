"""Build libnq_hip.so (HIP kernels + C ABI + C++ drop-in) and the `n_queens` CLI in-tree.

    python -m n_queens_amd.build            # both, for gfx950

Outputs: n_queens_amd/libnq_hip.so and n_queens_amd/n_queens (rpath $ORIGIN).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libnq_hip.so")
CLI = os.path.join(HERE, "n_queens")

ARCH = os.environ.get("NQ_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = os.environ.get("CXX", "g++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

# Packed-FP32 instructions are kept out of every code object (no float math here, but the
# flag is applied to all HIP code built by this project).
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + CSRC]

LIB_SOURCES = ["nq_dfs.hip", "nq_expand.hip", "nq_static.hip", "nq_api.cpp", "nq_prefix.cpp", "nq_compat.cpp", "nq_batch.cpp", "nq_serve.cpp",
               "nq_resume.cpp", "nq_cpu.cpp", "nq_tune.cpp"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = True) -> str:
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    for src in LIB_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, f"--offload-arch={ARCH}", *DEVICE_FLAGS, *COMMON, "-c", path, "-o", obj]
        else:
            # Host-only C++: the system compiler against the HIP runtime headers.
            cmd = [CXX, *COMMON, *HOST_FLAGS, "-pthread", "-c", path, "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli_tmp = CLI + ".tmp"
    _run([CXX, *COMMON, *HOST_FLAGS, "-o", cli_tmp, os.path.join(CSRC, "nq_cli.cpp"), "-L" + HERE, "-lnq_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(cli_tmp, CLI)
    return LIB


PROBES = os.path.join(ROOT, "tests", "probes")


def build_probes() -> None:
    """The stand-alone GPU probes under tests/probes (timing programs, not part of the library)."""
    for src in sorted(os.listdir(PROBES)):
        if src.endswith(".hip"):
            exe = os.path.join(PROBES, os.path.splitext(src)[0])
            _run([HIPCC, f"--offload-arch={ARCH}", *DEVICE_FLAGS, "-O3", "-std=c++17", "-o", exe + ".tmp",
                  os.path.join(PROBES, src)])
            os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    build()
    build_probes()
    sys.exit(0)
