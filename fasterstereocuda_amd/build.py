"""Builds the native libraries in-tree (no JIT cache, nothing installed).

  fasterstereocuda_amd/libfsgm.so              HIP kernels + C ABI (hipcc, gfx950)
  fasterstereocuda_amd/libStereoCuda.so        StereoCuda drop-in class over the C ABI
  fasterstereocuda_amd/libFasterStereoCuda.so  FasterStereoCuda drop-in class over the C ABI
  fasterstereocuda_amd/stereo_demo             C++ program that drives both classes
  fasterstereocuda_amd/stereo_console          the reference console's command line
                                               (option.xml, PNG / PGM in, PFM / PNG out)

Run as ``python -m fasterstereocuda_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["sgm_kernels.hip", "sgm_wide.hip", "sgm_capi.hip", "dsgm_kernels.hip",
               "dsgm_capi.hip", "pyr_kernels.hip", "pyr_capi.hip", "peaks.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    # no packed-FP32 VALU instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIB = os.path.join(PKG, "libfsgm.so")
CXX_LIB = os.path.join(PKG, "libStereoCuda.so")
FS_LIB = os.path.join(PKG, "libFasterStereoCuda.so")
DEMO = os.path.join(PKG, "stereo_demo")
CONSOLE = os.path.join(PKG, "stereo_console")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP toolchain is required to build libfsgm.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build failed ({res.returncode}):\n{' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_hip(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, "-shared", "-I", INCLUDE, "-I", CSRC,
           *[os.path.join(CSRC, s) for s in HIP_SOURCES], "-o", tmp]
    _run(cmd, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cxx(verbose: bool = False) -> None:
    cxx = shutil.which("g++") or "g++"
    rpath = "-Wl,-rpath,$ORIGIN"
    for out, src in ((CXX_LIB, "StereoCuda.cpp"), (FS_LIB, "FasterStereoCuda.cpp")):
        tmp = out + ".tmp"
        _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, os.path.join(CSRC, src),
              "-L", PKG, "-lfsgm", rpath, "-o", tmp], verbose)
        os.replace(tmp, out)
    tmp = DEMO + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-I", INCLUDE, os.path.join(CSRC, "stereo_demo.cpp"),
          "-L", PKG, "-lStereoCuda", "-lFasterStereoCuda", "-lfsgm", rpath, "-o", tmp], verbose)
    os.replace(tmp, DEMO)
    tmp = CONSOLE + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE,
          *[os.path.join(CSRC, f) for f in ("stereo_console.cpp", "option_xml.cpp", "image_io.cpp")],
          "-L", PKG, "-lFasterStereoCuda", "-lfsgm", "-lz", rpath, "-o", tmp], verbose)
    os.replace(tmp, CONSOLE)


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_cxx(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB, CXX_LIB, FS_LIB, DEMO, CONSOLE)
