"""Build the HIP engine (libibem3d.so) for gfx950 with hipcc, in-tree.

All kernels and the C-ABI are compiled in one hipcc invocation per translation unit, then linked
into ``ibemproject_amd/libibem3d.so``. Packed FP32 instructions are disabled
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libibem3d.so")
SOURCES = ["assembly.hip", "solver.hip", "post.hip", "api.cpp", "batch.cpp", "learned.hip"]
HOST_SOURCES = ["io.cpp"]           # plain C++ (no HIP), compiled with the host compiler
DRIVER = os.path.join(HERE, "ibem3d_run")
ARCH = os.environ.get("IBEM3D_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", CSRC, "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
        return obj

    def compile_host(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-6000:]}")
        return obj

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, SOURCES)) + list(ex.map(compile_host, HOST_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # the native driver links the library from its own directory
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-o", DRIVER + ".tmp",
           os.path.join(CSRC, "ibem3d_run.cpp"), "-L" + HERE, "-libem3d",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed:\n{r.stderr[-4000:]}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
