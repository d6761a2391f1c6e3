"""Build the HIP library fwi_amd/libfwi.so for gfx950 (and the C driver) with hipcc.

Flags that matter for correctness:
  -ffp-contract=off                         every stored value is the IEEE result of the written
                                            expression (no FMA contraction), as in the CPU oracle
  -fhip-fp32-correctly-rounded-divide-sqrt  float division correctly rounded
  -Xclang -target-feature -Xclang -packed-fp32-ops  no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfwi.so")
ARCH = os.environ.get("FWI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["legacy.hip", "plan.hip", "lbfgs.hip", "util.cpp"]
HEADERS = ["fwi_common.h", "fwi_internal.h", "batch_kernels.h", "batch_kernels.hip"]

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-command-line-argument",
]
DEVICE = [
    f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt", "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + ["-I", os.path.join(ROOT, "include")]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in (r.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if warn and os.environ.get("FWI_BUILD_VERBOSE"):
        print(warn, file=sys.stderr)
    return out


DRIVER_SRC = os.path.join(HERE, "driver", "fwi_main.c")
DRIVER_BIN = os.path.join(HERE, "fwi_main")


def build_driver() -> str:
    """The C driver (src/main.c's control flow on the C ABI), linked against libfwi.so."""
    cc = shutil.which("gcc") or "cc"
    tmp = DRIVER_BIN + ".tmp"
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-o", tmp, "-L", HERE, "-lfwi",
           "-L/opt/rocm/lib", "-lamdhip64", "-lm", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER_BIN)
    return DRIVER_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print(f"built {LIB} and {DRIVER_BIN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
