"""Build the HIP library (gfx950) and the C++ host programs in-tree.

    python -m seismic_tomography_3d_amd.build          # everything
    python -m seismic_tomography_3d_amd.build --hip    # libst_hip.so only

Outputs (git-ignored, they travel with the source tree):
    seismic_tomography_3d_amd/libst_hip.so      C ABI of include/st_api.h
    seismic_tomography_3d_amd/libst_host.so     C++ Eikonal/Tomography mirror
    seismic_tomography_3d_amd/bin/st_eikonal    eikonal_main equivalent
    seismic_tomography_3d_amd/bin/st_tomography tomography_main equivalent
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function", "-Wno-unused-variable",
    # packed FP32 VALU forms stay disabled (v_pk_* f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
# files whose results must be bit-identical to the CPU formulas
NO_CONTRACT = {"picks.hip", "tomo.hip", "fsm.hip", "wave.hip", "podvin.hip"}
# files whose kernels never form or receive a NaN (inputs are checked or
# sanitised on the bits); lets min/max skip quieting their operands
NO_NANS = {"fim.hip"}

HIP_SOURCES = ["api.hip", "fim.hip", "picks.hip", "tomo.hip", "surrogate.hip", "nn.hip", "fsm.hip",
               "wave.hip", "podvin.hip"]
HOST_SOURCES = ["file_manager.cpp", "geometry.cpp", "device_eikonal.cpp", "block_fim.cpp",
                "accurate_fsm.cpp", "classic.cpp", "least_squares.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r.stdout


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    flags = list(HIP_FLAGS)
    if os.path.basename(src) in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    if os.path.basename(src) in NO_NANS:
        flags.append("-fno-honor-nans")
    _run([HIPCC, *flags, "-c", src, "-o", obj])
    return obj


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    out = os.path.join(PKG, "libst_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host(jobs=8):
    """C++ mirror of the reference Eikonal/Tomography classes over the C ABI."""
    present = [s for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not present:
        return None
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    common = ["-std=c++17", "-O2", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{HOST}"]
    lib = os.path.join(PKG, "libst_host.so")
    rpath = "-Wl,-rpath,$ORIGIN"
    _run([cxx, *common, "-shared", *[os.path.join(HOST, s) for s in present],
          f"-L{PKG}", "-l:libst_hip.so", rpath, "-o", lib + ".tmp"])
    os.replace(lib + ".tmp", lib)
    for main, exe in (("eikonal_main.cpp", "st_eikonal"), ("tomography_main.cpp", "st_tomography")):
        src = os.path.join(HOST, main)
        if os.path.exists(src):
            out = os.path.join(PKG, "bin", exe)
            _run([cxx, *common, src, f"-L{PKG}", "-l:libst_host.so", "-l:libst_hip.so",
                  "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"])
            os.replace(out + ".tmp", out)
    return lib


def build_oracle():
    """The checker: oracle/liboracle.so and, where the reference checkout is
    present, oracle/_ref/libstref.so."""
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "all"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip", action="store_true")
    ap.add_argument("--host", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    a = ap.parse_args(argv)
    everything = not (a.hip or a.host or a.oracle)
    if everything or a.hip:
        print("built", build_hip())
    if everything or a.host:
        print("built", build_host())
    if everything or a.oracle:
        build_oracle()
        print("built oracle")


if __name__ == "__main__":
    main()
