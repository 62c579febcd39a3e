"""Build the HIP library (libbfpn_hip.so) and the bfp_solver CLI for gfx950.

Every source is compiled on each call (no dependency tracking), in parallel,
into bfpn_gpu_amd/_build/, then linked in-tree so that the shared library and
the executable travel with the source tree.

Flags:
  --offload-arch=gfx950                       MI355X (CDNA4) code objects
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                              no packed FP32 instructions
  -mllvm -amdgpu-mfma-vgpr-form              MFMA results in ordinary VGPRs
                                              (no copies out of AGPRs)
  -fvisibility=hidden                         only the BFPN_API entry points
                                              of include/bfpn.h are exported
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbfpn_hip.so")
CLI = os.path.join(PKG, "bin", "bfp_solver")

ARCH = os.environ.get("BFPN_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

LIB_SOURCES = ["tables.cpp", "energy.hip", "source.hip", "transport.hip", "angle.hip", "misc.hip",
               "diag.hip", "solver.hip", "server.hip"]
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result",
          f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form", "-fvisibility=hidden", "-fno-gpu-rdc"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", "-o", obj, path]
    else:  # host-only C++ (tables): same compiler, no device code
        cmd = [HIPCC, *COMMON, "-fvisibility=hidden", "-c", "-o", obj, path]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    main_obj = os.path.join(BUILD, "bfp_solver_main.o")
    _run([HIPCC, *COMMON, "-c", "-o", main_obj, os.path.join(CSRC, "bfp_solver_main.cpp")])
    _run([HIPCC, "-o", CLI, main_obj, f"-L{PKG}", "-lbfpn_hip", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
