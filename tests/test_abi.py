"""The C ABI (include/*.h) on the CPU: every declared function is exported by
the built libraries and bound by the ctypes layer, the record layouts the
ctypes layer uses match the C compiler's, and status codes / version are as
declared.  No GPU calls."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from j3d_amd import ffi, meshio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "j3d_amd", "lib", "libj3dg.so")


def declared_functions(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(j3dg_[a-z0-9_]+)\s*\(", text)) - set(re.findall(r"#define\s+(j3dg_\w+)", text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.mark.parametrize("header", ["j3dg.h", "j3dg_host.h", "j3dg_scenegen.h"])
def test_every_declared_function_is_exported(header):
    names = declared_functions(header)
    assert names, header
    exp = exported(LIB)
    missing = [n for n in names if n not in exp]
    assert not missing, missing


def test_io_functions_are_exported_by_the_canvas_library():
    """include/j3dg_io.h is the canvas layer's (libj3d_canvas.so), bound by j3d_amd.meshio."""
    from j3d_amd import meshio
    names = declared_functions("j3dg_io.h")
    assert names
    exp = exported(meshio.CANVAS_LIB)
    assert not [n for n in names if n not in exp]
    lib = meshio._load()
    for n in names:
        assert getattr(lib, n).argtypes is not None or getattr(lib, n).restype is not None, n


def test_ctypes_binds_every_function():
    lib = ffi.load()
    names = set()
    for h in ("j3dg.h", "j3dg_host.h", "j3dg_scenegen.h"):
        names |= set(declared_functions(h))
    for n in names:
        assert hasattr(lib, n), n
    unbound = sorted(n for n in names if n not in ffi.SIGNATURES)
    assert not unbound, unbound


def test_record_layouts_match_the_c_compiler(tmp_path):
    structs = {"j3dg_pixel": None, "j3dg_frame_desc": ffi.FrameDesc, "j3dg_build_opts": ffi.BuildOpts,
               "j3dg_mesh_info": ffi.MeshInfo, "j3dg_tuning": ffi.Tuning, "j3dg_stats": ffi.Stats,
               "j3dg_train_opts": ffi.TrainOpts, "j3dg_train_report": ffi.TrainReport,
               "j3dg_autotune_report": ffi.AutotuneReport, "j3dg_camera": ffi.Camera,
               "j3dg_mesh_arrays": meshio.MeshArrays}
    src = tmp_path / "sizes.c"
    body = "\n".join('printf("%s %%zu\\n", sizeof(%s));' % (s, s) for s in structs)
    src.write_text('#include <stdio.h>\n#include "j3dg.h"\n#include "j3dg_host.h"\n#include "j3dg_io.h"\n'
                   'int main(void){%s return 0;}\n' % body)
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I" + INC, "-o", str(exe), str(src)], check=True)
    sizes = dict(l.split() for l in subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                                                   check=True).stdout.splitlines())
    assert int(sizes["j3dg_pixel"]) == ffi.PIXEL_DTYPE.itemsize == 32
    for name, cls in structs.items():
        if cls is not None:
            assert int(sizes[name]) == C.sizeof(cls), name


def test_version_and_status_codes():
    lib = ffi.load()
    assert lib.j3dg_abi_version() == 2
    text = open(os.path.join(INC, "j3dg.h")).read()
    for code, name in ffi.ERRORS.items():
        assert re.search(r"#define J3DG_ERR_\w+ %d\b" % code, text), (code, name)


def test_null_arguments_are_rejected_without_a_device():
    lib = ffi.load()
    assert lib.j3dg_create(0, None) == 1          # invalid argument
    assert lib.j3dg_device_count(None) == 1
    assert lib.j3dg_set_tuning(None, None) == 1   # no context
    assert lib.j3dg_default_train_opts(None) == 1
    o = ffi.TrainOpts()
    assert lib.j3dg_default_train_opts(C.byref(o)) == 0
    assert o.n_rays == 1 << 24 and o.batch % 128 == 0


def test_frame_object_list():
    """More objects than j3dg_frame_desc::objects holds go by object_list,
    which the descriptor keeps alive; up to 16 fill objects[]."""
    import j3d_amd as J
    cam = J.camera_with_vertical_fov(60.0)
    cs, ci = J.translation_camera(0.0, 0.0, 3.0)
    small = J.make_frame(8, 8, cam, cs, ci, np.zeros(3, np.float32), 1.0, list(range(16))).desc()
    assert small.n_objects == 16 and not small.object_list and list(small.objects) == list(range(16))
    slots = [3 * i + 1 for i in range(40)]
    big = J.make_frame(8, 8, cam, cs, ci, np.zeros(3, np.float32), 1.0, slots).desc()
    assert big.n_objects == 40 and big.object_list
    got = np.ctypeslib.as_array((C.c_uint32 * 40).from_address(big.object_list))
    assert list(got) == slots


def test_point_cloud_arguments_are_checked_before_the_c_call():
    """Mismatched per-point arrays and matrices that are not 16 floats raise
    ValueError in the Python layer (the C side copies n * 12 bytes of normals
    and reads 16 floats per matrix); no context or GPU is involved."""
    import j3d_amd as J
    r = J.Renderer.__new__(J.Renderer)  # the checks run before any library call
    r.ctx = None
    p = np.zeros((10, 3), np.float32)
    with pytest.raises(ValueError):
        r.upload_points(p, normals=np.zeros((9, 3), np.float32))
    with pytest.raises(ValueError):
        r.upload_points(p, colors=np.zeros(11, np.uint32))
    with pytest.raises(ValueError):
        r.upload_points(p, cs=np.eye(3, dtype=np.float32))
    with pytest.raises(ValueError):
        r.set_points_cs(0, np.zeros(12, np.float32))
    with pytest.raises(ValueError):
        r.splat_points(0, np.zeros(16, np.float32), np.zeros(15, np.float32), 0, [0])
    with pytest.raises(ValueError):
        r.splat_points(0, np.zeros(9, np.float32), np.zeros(16, np.float32), 0, [0])


def test_build_lists_only_existing_sources():
    from j3d_amd import build as B
    for s in B.HIP_SOURCES + B.CPP_SOURCES:
        assert os.path.exists(os.path.join(B.CSRC, s)), s


def test_proposal_checkpoint_round_trip_and_refusals(tmp_path):
    """The proposal network's checkpoint (magic, version, layer sizes,
    normalisation, parameters) reads back bit-for-bit; wrong parameter or
    normalisation counts raise ValueError before anything is written, and a
    foreign, truncated or re-versioned file is refused on reading."""
    from j3d_amd import api
    rng = np.random.default_rng(7)
    p = rng.standard_normal(ffi.MLP_PARAMS).astype(np.float32)
    nrm = np.array([0.5, -1.0, 2.0, 3.5], np.float32)
    path = str(tmp_path / "net.bin")
    api.write_proposal(path, p, nrm)
    p2, nrm2 = api.read_proposal(path)
    assert np.array_equal(p2, p) and np.array_equal(nrm2, nrm)
    with pytest.raises(ValueError):
        api.write_proposal(str(tmp_path / "short.bin"), p[:-1], nrm)
    with pytest.raises(ValueError):
        api.write_proposal(str(tmp_path / "norm.bin"), p, nrm[:3])
    assert not os.path.exists(str(tmp_path / "short.bin"))
    data = open(path, "rb").read()
    for name, blob in (("magic", b"X" + data[1:]), ("cut", data[:-4]), ("long", data + b"\0" * 4),
                       ("version", data[:8] + np.uint32(api.PROPOSAL_VERSION + 1).tobytes() + data[12:])):
        bad = str(tmp_path / (name + ".bin"))
        with open(bad, "wb") as f:
            f.write(blob)
        with pytest.raises(ValueError):
            api.read_proposal(bad)
    import j3d_amd as J
    r = J.Renderer.__new__(J.Renderer)  # the check runs before any library call
    r.ctx = None
    with pytest.raises(ValueError):
        r.set_proposal(0, p[:10], nrm)
