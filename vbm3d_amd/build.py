"""Builds libvbm3d_hip.so (gfx950) in-tree with hipcc.

Flags that matter for results:
  -ffp-contract=off                        no FMA contraction: the kernels
                                           reproduce the reference's
                                           separately rounded mul and add;
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE float division and sqrt;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvbm3d_hip.so")
# one translation unit per kernel family, compiled in parallel
SOURCES = [os.path.join(CSRC, f) for f in ("kernels_bm.hip", "kernels_filter.hip", "kernels_agg.hip",
                                           "kernels_gen.hip", "flow.hip", "xform_bench.hip", "pyramid.hip", "capi.hip")]
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VBM3D_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Werror=uninitialized",
    f"-I{os.path.join(ROOT, 'include')}",
]


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libvbm3d_host.so")
CLI = os.path.join(HERE, "bin", "VBM3Ddenoising")
FLOW_CLI = os.path.join(os.path.dirname(CLI), "tvl1flow")
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}"]
IMG_INC = os.environ.get("VBM3D_IMAGE_INCLUDE", "/opt/conda/include")  # png.h, jpeglib.h


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = False) -> str:
    """Compile libvbm3d_hip.so (HIP, gfx950), then the C++ host library
    libvbm3d_host.so (run_vbm3d over the C ABI) and the VBM3Ddenoising
    command line, both plain host C++ linked against libvbm3d_hip.so."""
    os.makedirs(OBJDIR, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        cmd = [HIPCC, *FLAGS, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    tmp = LIB + ".tmp"
    # rocBLAS: the pyramid's full-frame DCTs as DGEMMs
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocblas",
          "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    host_src = [os.path.join(HOST, "vbm3d_host.cpp"), os.path.join(HOST, "image_io.cpp")]
    # PNG through the system's libpng16 (headers of the same 1.6 release
    # from the image's conda tree, searched after the system's); JPEG's
    # libjpeg 9 is loaded at run time by image_io.cpp
    _run([CXX, *CXXFLAGS, "-idirafter", IMG_INC, "-fPIC", "-shared", "-o", HOST_LIB + ".tmp", *host_src,
          f"-L{HERE}", "-l:libvbm3d_hip.so", "-l:libpng16.so.16", "-ldl", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    _run([CXX, *CXXFLAGS, "-o", CLI + ".tmp", os.path.join(HOST, "main.cpp"),
          f"-L{HERE}", "-l:libvbm3d_host.so", "-l:libvbm3d_hip.so", "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(CLI + ".tmp", CLI)
    # the TV-L1 flow command line (the reference's tvl1flow program)
    _run([CXX, *CXXFLAGS, "-o", FLOW_CLI + ".tmp", os.path.join(HOST, "tvl1flow_main.cpp"),
          f"-L{HERE}", "-l:libvbm3d_host.so", "-l:libvbm3d_hip.so", "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(FLOW_CLI + ".tmp", FLOW_CLI)
    # the multiscale pyramid programs (the reference's decompose, recompose, merge_coarse)
    for tool, name in enumerate(("decompose", "recompose", "merge_coarse")):
        exe = os.path.join(os.path.dirname(CLI), name)
        _run([CXX, *CXXFLAGS, f"-DPYR_TOOL={tool}", "-o", exe + ".tmp", os.path.join(HOST, "pyramid_main.cpp"),
              f"-L{HERE}", "-l:libvbm3d_host.so", "-l:libvbm3d_hip.so", "-Wl,-rpath,$ORIGIN/.."], verbose)
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
