"""Build the native parts of diff_tritracer in-tree.

  libtritracer_hip.so : HIP kernels (gfx950) + the C ABI of include/tritracer_hip.h  (hipcc)
  _C.so               : pybind11 module binding OptixTracer (tracer.h) over that ABI  (g++, torch headers)

Both land in diff_tritracer/ next to __init__.py; _C.so finds the kernel library
through an $ORIGIN rpath. Run:  python diff-triangle-tracer_amd/build_native.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "diff_tritracer")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIP_LIB = os.path.join(PKG, "libtritracer_hip.so")
EXT_LIB = os.path.join(PKG, "_C.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["tt_bvh.hip", "tt_trace.hip", "tt_multi.hip", "tt_proposal.hip", "tt_camera.hip", "tt_optim.hip",
               "tt_densify.hip", "tt_sceneio.hip", "tt_capi.cpp"]
EXT_SOURCES = ["ext.cpp", "tracer.cpp"]

# -ffp-contract=off: the kernels round where the CPU oracle rounds (bit-exact hit lists).
# -packed-fp32-ops disabled: packed FP32 forms are avoided on this hardware.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", f"-I{INCLUDE}",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_hip(verbose: bool = False) -> str:
    """Compile the kernel library with hipcc for gfx950 (one object per source, in parallel)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("hipcc failed:\n  " + " ".join(cmd) + "\n" + out + err)
        if verbose and (out or err):
            print(out + err)
    tmp = HIP_LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_ext(verbose: bool = False) -> str:
    """Compile the pybind11 module _C (host C++ only) against torch and the kernel library."""
    import torch
    from torch.utils import cpp_extension

    incs = cpp_extension.include_paths() + [sysconfig.get_paths()["include"], os.path.join(ROCM, "include"), INCLUDE]
    libdirs = cpp_extension.library_paths()
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
           "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM", "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}"]
    cmd += [f"-I{i}" for i in incs]
    cmd += [os.path.join(CSRC, s) for s in EXT_SOURCES]
    cmd += [f"-L{d}" for d in libdirs] + [f"-L{PKG}", "-L" + os.path.join(ROCM, "lib")]
    cmd += ["-ltritracer_hip", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
            "-lamdhip64", "-Wl,-rpath,$ORIGIN"]
    tmp = EXT_LIB + ".tmp"
    cmd += ["-o", tmp]
    r = _run(cmd)
    if verbose and (r.stdout or r.stderr):
        print(r.stdout + r.stderr)
    os.replace(tmp, EXT_LIB)
    return EXT_LIB


def build_all(verbose: bool = False):
    return build_hip(verbose), build_ext(verbose)


if __name__ == "__main__":
    for p in build_all(verbose="-v" in sys.argv):
        print("built", p)
