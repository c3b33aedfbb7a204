"""Builds emu_amd/libemu_amd.so (HIP kernels + C ABI) for gfx950 with hipcc,
and the C++ host layer over it: emu_amd/libemu_cpp.so (cpp/Emu.hpp, the
reference's call surface) and emu_amd/bin/emu_main (the reference's
executable).

    python -m emu_amd.build        # or emu_amd.build.build()

Every source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).  emu_ref.hip, the
per-function entry points with the reference's rounding, is also built with
-ffp-contract=off.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libemu_amd.so")
OUT_CPP = os.path.join(HERE, "libemu_cpp.so")
OUT_MAIN = os.path.join(HERE, "bin", "emu_main")
CPP = os.path.join(HERE, "cpp")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__",
            "-I", os.path.join(ROOT, "include"), "-I", CPP, "-I", os.path.join(ROCM, "include")]
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("EMU_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]
SOURCES = {
    "emu_ref.hip": ["-ffp-contract=off"],
    "emu_cells.hip": [],
    "emu_rk.hip": ["-ffp-contract=off"],
    "emu_state.hip": [],
    "emu_init.hip": [],
    # MFMA accumulators in VGPRs: in the adjoint deposit's per-key loop the
    # default AGPR form copies them out and back at every k-step
    "emu_adjoint.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []
    procs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [cc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = OUT + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, OUT)
    build_cpp(verbose)
    return OUT


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed:\n{r.stdout}")


def build_cpp(verbose: bool = False):
    """The C++ host layer: host code only (g++), linked to libemu_amd.so."""
    cxx = os.environ.get("CXX", "g++")
    os.makedirs(os.path.dirname(OUT_MAIN), exist_ok=True)
    objs = []
    for src in ("Emu.cpp", "PlotFile.cpp", "emu_cpp.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run([cxx, *CXXFLAGS, "-c", os.path.join(CPP, src), "-o", obj], verbose)
        objs.append(obj)
    rocm_lib = os.path.join(ROCM, "lib")
    tmp = OUT_CPP + ".tmp"
    _run([cxx, "-shared", "-o", tmp, *objs, "-L", HERE, "-lemu_amd", "-L", rocm_lib, "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + rocm_lib], verbose)
    os.replace(tmp, OUT_CPP)
    tmp = OUT_MAIN + ".tmp"
    _run([cxx, *CXXFLAGS, os.path.join(CPP, "main.cpp"), "-o", tmp, "-L", HERE, "-lemu_cpp", "-lemu_amd",
          "-L", rocm_lib, "-lamdhip64", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + rocm_lib], verbose)
    os.replace(tmp, OUT_MAIN)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
