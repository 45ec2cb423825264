"""Build recipe for the gfx950 HIP library ``libtse_hip.so`` (in-tree).

``python -m transportse_amd.build`` compiles every ``csrc/*.hip`` translation
unit with hipcc for gfx950 and links them into
``transportse_amd/libtse_hip.so``, which the ctypes loader in ``_lib.py``
opens. The faithful kernels are compiled with ``-ffp-contract=off`` so that
no multiply-add is fused (reference arithmetic); the performance kernels may
contract. Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libtse_hip.so")
BUILD_DIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("TSE_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", INCLUDE, "-Wall", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable"]

# translation unit -> extra flags
UNITS = {
    "tse_core.hip": ["-ffp-contract=off"],
    "tse_faithful.hip": ["-ffp-contract=off"],
    "tse_fast.hip": ["-ffp-contract=fast"],
    "tse_batch.hip": ["-ffp-contract=fast"],
    "tse_step.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit: str, flags: list[str]) -> str:
    src = os.path.join(CSRC, unit)
    obj = os.path.join(BUILD_DIR, unit.replace(".hip", ".o"))
    cmd = [hipcc(), *COMMON, *flags, "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{proc.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile all units and link libtse_hip.so; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(UNITS))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
           "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_c_example()
    if verbose:
        print("built", LIB)
    return LIB


EXAMPLE_SRC = os.path.join(os.path.dirname(HERE), "tests", "c_abi_example.c")
EXAMPLE_BIN = os.path.join(HERE, "c_abi_example")


def build_c_example(out: str = EXAMPLE_BIN) -> str:
    """Compile the plain-C host of the C-ABI (tests/c_abi_example.c) with gcc
    as C99 against include/tse_hip.h and libtse_hip.so."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror",
           "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE, "-I", os.path.join(rocm, "include"),
           EXAMPLE_SRC, "-L", HERE, "-ltse_hip", "-L", os.path.join(rocm, "lib"),
           "-lamdhip64", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{os.path.join(rocm, 'lib')}",
           "-lm", "-o", out]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"C example failed to build:\n{proc.stderr[-4000:]}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
