"""Build recipe for the gfx950 SpTRSV library (in-tree, no JIT cache).

Outputs (next to this file):
  libcapellini_sptrsv.so   the C-ABI library (HIP kernels + host code)
  sptrsv_main              drop-in driver, VALUE_TYPE=double
  sptrsv_main_f32          drop-in driver, VALUE_TYPE=float
and, under oracle/, the test-only CPU oracle (liboracle.so) plus, when the
reference tree is present, oracle/_ref/ (the reference's own host helpers).

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled and FP contraction off (the kernels' fused multiply-adds
are explicit, so the arithmetic is exactly what the source says).
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
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcapellini_sptrsv.so")
MAIN = os.path.join(PKG, "sptrsv_main")
MAIN32 = os.path.join(PKG, "sptrsv_main_f32")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-ffp-contract=off"]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out)


def build_library(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = {
        "host_prep.o": [CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "host_prep.cpp"), "-o",
                        os.path.join(BUILD, "host_prep.o")],
        "solve.o": [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "solve.hip"), "-o",
                    os.path.join(BUILD, "solve.o")],
    }
    with ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(_run, objs.values()))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp,
          *[os.path.join(BUILD, o) for o in objs],
          f"-Wl,--version-script={os.path.join(CSRC, 'exports.map')}"])
    os.replace(tmp, LIB)
    for out, vt in ((MAIN, "double"), (MAIN32, "float")):
        _run([CXX, *CXX_FLAGS, f"-DVALUE_TYPE={vt}", os.path.join(CSRC, "main.cpp"), "-o", out,
              f"-L{PKG}", "-lcapellini_sptrsv", "-Wl,-rpath,$ORIGIN",
              "-L/opt/rocm/lib", "-lamdhip64"])
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = True) -> None:
    oracle_dir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle_dir, "oracle"])
    ref = os.environ.get("CAPELLINI_REFERENCE", "/root/reference")
    if os.path.isdir(ref) and shutil.which("g++"):
        _run(["make", "-s", "-B", "-C", oracle_dir, "ref", f"REF={ref}"])
        if verbose:
            print("built oracle/_ref from the reference tree")


def build(verbose: bool = True) -> None:
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
