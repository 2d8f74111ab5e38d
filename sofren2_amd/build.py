"""Builds libsofren2_gpu.so (HIP backend for gfx950) and libsofren2_scenes.so.

Host code that computes floats fed to the GPU (frustum planes, dmesh vertices,
mip chains) is compiled with -ffp-contract=off -fno-builtin so that its results
match the reference build bit for bit; device code with -ffp-contract=off so
every FMA is an explicit fmaf.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sofren2_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_FLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-fno-builtin", "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xarch_host", "-fno-builtin",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

LIBNAME = os.path.join(LIB, "libsofren2_gpu.so")
SCENES = os.path.join(LIB, "libsofren2_scenes.so")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in ("frontend.cpp", "host_queries.cpp", "builders.cpp", "rcp_probe.cpp", "tlas.cpp", "server.cpp"):
        o = os.path.join(OBJ, src + ".o")
        _run(["g++", "-std=c++17", "-c"] + HOST_FLAGS + [os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in ("backend.cpp", "kernels.hip"):
        o = os.path.join(OBJ, src + ".o")
        _run([HIPCC, "-x", "hip", "-c"] + HIP_FLAGS + [os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIBNAME + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-Wl,-Bsymbolic"])
    os.replace(tmp, LIBNAME)
    tmp = SCENES + ".tmp"
    _run(["gcc", "-std=gnu11", "-shared"] + HOST_FLAGS + [
        "-I" + os.path.join(PKG, "scenes"), os.path.join(PKG, "scenes", "scenes.c"),
        "-o", tmp, "-L" + LIB, "-lsofren2_gpu", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, SCENES)
    if verbose:
        print("built", LIBNAME, SCENES)
    return LIBNAME


if __name__ == "__main__":
    build(verbose=True)
