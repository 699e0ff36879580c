"""Build recipe of libsoapgpu.so (gfx950) and the C++ command-line tools.

hipcc compiles the HIP sources for gfx950 only; the library is written in-tree
(soap_gpu_amd/libsoapgpu.so) so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsoapgpu.so")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
]
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("soap_kernels.hip", "soap_linear.hip", "soap_diff.hip", "soap_inject.hip",
                                                          "soap_phoenix.hip", "soap_engine.hip")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_lib(verbose: bool = False, out: str = LIB, defines=()) -> str:
    objs = []
    bdir = os.path.join(HERE, "build", os.path.basename(out) + ".d")
    os.makedirs(bdir, exist_ok=True)
    for s in SOURCES:
        o = os.path.join(bdir, os.path.basename(s) + ".o")
        _run([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-I", os.path.join(ROOT, "include"), "-c", s, "-o", o])
        objs.append(o)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


# executables keep the names of the reference's stages (its os.system calls)
TOOLS = {
    "SOAP_initialize": ("cli_initialize.cpp", []),
    "SOAP_integration": ("cli_integration.cpp", []),
    "SOAP_integration_solar": ("cli_integration.cpp", ["-DSOAP_SOLAR_MODE=1"]),
    "SOAP_resolution": ("cli_resolution.cpp", []),
    "BIS_FIT2": ("cli_bisfit.cpp", []),
    "soap_run": ("soap_run.cpp", []),
    "phoenix_extraction": ("cli_phoenix.cpp", []),
}


def build_tools(verbose: bool = False) -> list[str]:
    """C++ host driver and CLI shims (g++, linked against libsoapgpu.so)."""
    drv = os.path.join(HERE, "driver")
    os.makedirs(BIN, exist_ok=True)
    out = []
    for exe_name, (src, extra) in TOOLS.items():
        exe = os.path.join(BIN, exe_name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", *extra, "-I", os.path.join(ROOT, "include"), "-I", drv,
              os.path.join(drv, src), "-o", exe, "-L", HERE, "-lsoapgpu",
              "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + HERE])
        out.append(exe)
        if verbose:
            print("built", exe)
    return out


def build(verbose: bool = False):
    build_lib(verbose)
    build_tools(verbose)


if __name__ == "__main__":
    build(verbose=True)
