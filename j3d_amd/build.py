"""Builds the gfx950 library in-tree: j3d_amd/lib/libj3dg.so.

HIP sources are compiled by hipcc for gfx950 only; host-only C++ sources by
g++.  Every translation unit uses -ffp-contract=off (no fused multiply-add),
keeps fp32 denormals and correctly rounded division/sqrt, so the kernels
round exactly like the host-side arithmetic.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "j3d_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-DNDEBUG",
          "-I" + INCLUDE, "-I" + CSRC]
HIP_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-fno-gpu-flush-denormals-to-zero",
             "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]

HIP_SOURCES = ["j3dg.hip", "mlp.hip", "bvh_gpu.hip", "allhits.hip", "splat.hip"]
CPP_SOURCES = ["bvh_build.cpp", "scenegen.cpp", "host_tables.cpp", "autotune.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r.stdout


def _compile_all(jobs):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), 8))) as ex:
        list(ex.map(_run, jobs))


def build_library(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs, jobs = [], []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise RuntimeError("build: listed source %s is missing" % path)
        obj = os.path.join(OBJDIR, src + ".o")
        jobs.append([HIPCC] + COMMON + HIP_FLAGS + ["-c", "-o", obj, path])
        objs.append(obj)
    for src in CPP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            raise RuntimeError("build: listed source %s is missing" % path)
        obj = os.path.join(OBJDIR, src + ".o")
        jobs.append([CXX] + COMMON + ["-pthread", "-c", "-o", obj, path])
        objs.append(obj)
    _compile_all(jobs)
    lib = os.path.join(LIBDIR, "libj3dg.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build_canvas(verbose=False):
    """The C++ canvas/scene/camera layer over the C ABI and its headless harness."""
    src = os.path.join(HOST, "canvas.cpp")
    if not os.path.exists(src):
        return None
    lib = os.path.join(LIBDIR, "libj3d_canvas.so")
    _run([CXX] + COMMON + ["-pthread", "-shared", "-I" + HOST, "-o", lib, src, os.path.join(HOST, "ply.cpp"),
                           "-L" + LIBDIR, "-lj3dg", "-Wl,-rpath,$ORIGIN"])
    demo = os.path.join(HOST, "render_frame.cpp")
    if os.path.exists(demo):
        exe = os.path.join(LIBDIR, "j3d_render_frame")
        _run([CXX] + COMMON + ["-pthread", "-I" + HOST, "-o", exe, demo,
                               "-L" + LIBDIR, "-lj3d_canvas", "-lj3dg", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", lib)
    return lib


def build_oracle(verbose=False):
    """CPU oracle (test infrastructure): oracle/liboracle.so."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built oracle/liboracle.so")


def build_all(verbose=False):
    build_library(verbose)
    build_canvas(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
