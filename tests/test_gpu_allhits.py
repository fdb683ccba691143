"""GPU parity of the all-hits query (jtk::qbvh::find_all_triangles, called at
j3d/vox.cpp:333) and of write_vox's voxel grid (j3d/vox.cpp:278-376) against
the oracle: hit lists and voxel grids bit-exact."""
import numpy as np
import pytest

import j3d_amd as J
import oracle as O
from j3d_amd import ffi
from scenes import random_rays

pytestmark = pytest.mark.gpu


def _cube_mesh():
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    q = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    t = np.array([[a, b, c] for a, b, c, d in q] + [[a, c, d] for a, b, c, d in q], np.uint32)
    return v, t


def _slabs(n):
    """n parallel unit squares x = 0.02 k: every x ray through them hits n
    triangles, more than the voxel kernel's first list size (32)."""
    v, t = [], []
    for k in range(n):
        b = len(v)
        x = np.float32(0.02) * np.float32(k)
        v += [[x, 0, 0], [x, 1, 0], [x, 1, 1], [x, 0, 1]]
        t += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    return np.array(v, np.float32), np.array(t, np.uint32)


def _compare_lists(got, ref, label):
    assert np.array_equal(got["count"], ref["count"]), label
    assert np.array_equal(got["tri"], ref["tri"]), label
    for k in ("t", "u", "v"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (label, k)


@pytest.mark.parametrize("width, builder", [(8, ffi.BUILD_HOST), (4, ffi.BUILD_HOST), (8, ffi.BUILD_DEVICE)])
def test_all_hits_matches_oracle(renderer, width, builder):
    v, t = J.torus(200, 90, seed=4)
    cs = np.eye(4, dtype=np.float32)
    cs[3, :3] = (0.25, -0.5, 0.125)  # column-major translation: world rays go to the mesh's frame
    cs = cs.reshape(16)
    slot = renderer.upload_mesh(v, t, cs=cs, width=width, builder=builder)
    m = O.Mesh(v, t, cs=cs)
    rays = random_rays(40000, 21, np.array([0.25, -0.5, 0.125]), 1.4)
    try:
        for max_hits in (16, 3, 0):
            got = renderer.all_hits(slot, rays, max_hits)
            ref = O.all_hits(m, rays, max_hits)
            _compare_lists(got, ref, "W%d max_hits %d" % (width, max_hits))
        print("W%d builder %d: %d rays, %d with >= 2 hits, at most %d" % (
            width, builder, len(rays), int((ref["count"] >= 2).sum()), int(ref["count"].max())))
        assert ref["count"].max() >= 4
    finally:
        renderer.remove_mesh(slot)


def test_voxelize_cube_shell(renderer):
    v, t = _cube_mesh()
    slot = renderer.upload_mesh(v, t)
    try:
        dims, data = renderer.voxelize(slot, 8)
        assert dims == (8, 8, 8)
        shell = np.ones((8, 8, 8), bool)
        shell[1:-1, 1:-1, 1:-1] = False
        assert np.array_equal(data > 0, shell)
        assert set(np.unique(data[shell]).tolist()) == {255}
    finally:
        renderer.remove_mesh(slot)


@pytest.mark.parametrize("source", ["white", "vertex", "texture"])
def test_voxelize_matches_oracle(renderer, source):
    v, t = J.torus(160, 80, seed=6)
    rng = np.random.default_rng(3)
    vc = uv = tex = None
    if source in ("vertex", "texture"):
        vc = rng.uniform(0, 1, size=(len(v), 3)).astype(np.float32)
    if source == "texture":
        uv = rng.uniform(-0.1, 1.1, size=(len(t), 6)).astype(np.float32)
        tex = rng.integers(0, 2 ** 32, size=(37, 53), dtype=np.uint64).astype(np.uint32)
    slot = renderer.upload_mesh(v, t)
    try:
        if vc is not None:
            renderer.set_mesh_colors(slot, vertex_colors=vc, uv=uv, texture=tex)
        dims, data = renderer.voxelize(slot, 64)
        rdims, ref = O.voxelize(O.Mesh(v, t, vertex_colors=vc, uv=uv, texture=tex), 64)
        assert dims == rdims
        assert np.array_equal(data, ref), int((data != ref).sum())
        print("%s: dims %s, %d voxels set, %d colours" % (source, dims, int((data > 0).sum()),
                                                            len(np.unique(data))))
        assert (data > 0).sum() > 1000
    finally:
        renderer.remove_mesh(slot)


def test_voxelize_long_hit_lists_and_device_build(renderer):
    """Rays with more hits than the first list size: the kernel reports it and
    the axis runs again with longer lists."""
    v, t = _slabs(50)
    slot = renderer.upload_mesh(v, t, builder=ffi.BUILD_DEVICE)
    try:
        dims, data = renderer.voxelize(slot, 40)
        rdims, ref = O.voxelize(O.Mesh(v, t), 40)
        assert dims == rdims and np.array_equal(data, ref)
        assert (data > 0).sum() >= 50 * 20
    finally:
        renderer.remove_mesh(slot)


def test_voxelize_mesh_box_not_finite(renderer):
    """A NaN first vertex makes that axis of the mesh box NaN (the box is
    std::min / std::max folded from the first vertex): one cell on that axis,
    every voxel ray starts at a NaN coordinate and hits nothing."""
    v, t = J.torus(30, 15, seed=2)
    v = v.copy()
    v[0, 1] = np.nan
    slot = renderer.upload_mesh(v, t)
    try:
        dims, data = renderer.voxelize(slot, 32)
        rdims, ref = O.voxelize(O.Mesh(v, t), 32)
        assert dims == rdims and dims[1] == 1 and np.array_equal(data, ref) and not data.any()
        rays = random_rays(500, 3, np.zeros(3), 1.4)
        got = renderer.all_hits(slot, rays, 4)
        assert np.array_equal(got["count"], O.all_hits(O.Mesh(v, t), rays, 4, brute=True)["count"])
    finally:
        renderer.remove_mesh(slot)
