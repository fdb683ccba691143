"""Builds the CPU emulation of libj3dg (test infrastructure): j3d_amd/csrc/j3dg.hip
compiled for x86 against tests/emu/hip/hip_runtime.h, with the host-only sources
of the library, into OUT/libj3dg_emu.so.

The kernel source is copied with a few textual edits that only the CPU build
needs: address-space qualifiers removed (one address space on the CPU; the LDS
copy of a top level keeps a type of its own so its overloads stay distinct),
the dynamic LDS array taken from the emulator, trace kernels registered when
their address is taken (hipLaunchKernel finds them by address), and batching
switched on.  `variant` plants the two mistakes the batch path once had, to show
the emulator finds what faulted the GPU:
  "wrong_table": the batch launch reads the single-frame object table;
  "no_frame_handoff": the tail modes move a ray without its frame.
The allocation-free entry point's fixed array is placed in fenced memory too.
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
CSRC = os.path.join(ROOT, "j3d_amd", "csrc")
EMU = os.path.dirname(os.path.abspath(__file__))


def _sub(s, a, b, count=1):
    n = s.count(a)
    if n != count:
        raise RuntimeError("emu build: expected %d occurrence(s) of %r, found %d" % (count, a[:60], n))
    return s.replace(a, b)


def patched_sources(out, variant=None):
    os.makedirs(out, exist_ok=True)
    for name in ("traverse.h", "rt_math.h", "j3dg_internal.h", "bvh_build.h"):
        s = open(os.path.join(CSRC, name)).read()
        s = s.replace("__attribute__((address_space(1)))", "").replace("__attribute__((address_space(3)))", "")
        if name == "traverse.h":
            s = _sub(s, "typedef const  TopNode lds_top;",
                     "struct lds_top_node : TopNode {};\ntypedef const lds_top_node lds_top;")
        open(os.path.join(out, name), "w").write(s)
    s = open(os.path.join(CSRC, "j3dg.hip")).read()
    s = s.replace("__attribute__((address_space(1)))", "").replace("__attribute__((address_space(3)))", "")
    s = _sub(s, "extern __shared__ uint2 lds_stack[];", "uint2* lds_stack = static_cast<uint2*>(::emu::dyn_lds());")
    s = _sub(s, "return count ? (void*)&k_trace<W, D, true, MODE> : (void*)&k_trace<W, D, false, MODE>;",
             'return count ? ::emu::reg(&k_trace<W, D, true, MODE>, "k_trace") '
             ': ::emu::reg(&k_trace<W, D, false, MODE>, "k_trace");')
    s = s.replace("static constexpr bool kBatchEnabled = false;", "static constexpr bool kBatchEnabled = true;")
    # the fixed working array of the allocation-free entry point, fenced like device memory
    s = s.replace("static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];",
                  "static unsigned char (&scratch)[kNoAllocScratchBytes] = "
                  "*static_cast<unsigned char (*)[kNoAllocScratchBytes]>(::emu::device_array(kNoAllocScratchBytes));")
    if variant == "wrong_table":
        s = _sub(s, "const DevObject* objs = mode == kModeBatch ? d_bobjs : d_objs;", "const DevObject* objs = d_objs;")
    elif variant == "no_frame_handoff":
        s = _sub(s, """    if (M == kModeBatch) {  // the ray's frame travels with it
      shfl_from(fi, src);
      shfl_from(lnobj, src);
      shfl_from(lobjs, src);
    }
""", "")
    elif variant is not None:
        raise ValueError(variant)
    open(os.path.join(out, "j3dg_emu.cpp"), "w").write(s)
    return out


def compiler():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or ""):
        if c and os.path.exists(c):
            return c
    return None


def build(out, variant=None, jobs=4):
    cxx = compiler()
    if cxx is None:
        raise RuntimeError("emu build needs clang++ (ext_vector_type)")
    src = patched_sources(os.path.join(out, "src"), variant)
    flags = ["-std=c++17", "-O1", "-fno-inline", "-fno-omit-frame-pointer", "-g", "-fPIC", "-ffp-contract=off",
             "-fno-fast-math", "-w", "-I" + os.path.join(EMU), "-I" + src, "-I" + os.path.join(ROOT, "include"),
             "-I" + CSRC, "-pthread"]
    units = [os.path.join(src, "j3dg_emu.cpp"), os.path.join(EMU, "emu_runtime.cpp"), os.path.join(EMU, "emu_stubs.cpp")]
    units += [os.path.join(CSRC, n) for n in ("bvh_build.cpp", "scenegen.cpp", "host_tables.cpp", "autotune.cpp")]
    objs, procs = [], []
    for u in units:
        o = os.path.join(out, os.path.basename(u) + ".o")
        objs.append(o)
        procs.append(subprocess.Popen([cxx] + flags + ["-c", "-o", o, u], stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    for p, u in zip(procs, units):
        log = p.communicate()[0]
        if p.returncode != 0:
            raise RuntimeError("emu build failed on %s:\n%s" % (u, log[-6000:]))
    lib = os.path.join(out, "libj3dg_emu.so")
    r = subprocess.run([cxx, "-shared", "-o", lib] + objs + ["-pthread", "-Wl,--no-undefined"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("emu link failed:\n" + r.stdout[-6000:])
    return lib


if __name__ == "__main__":
    print(build(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else None))
