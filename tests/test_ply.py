"""PLY reading and writing of the canvas layer (read_ply / write_ply, j3d/io.cpp:733-741;
include/j3dg_io.h).  The reference forwards both to jtk, which is absent from
the tree, so these cases are pinned by the published format itself: files
written byte by byte here in each of the three encodings, with the arrays
they must give worked out by hand (fan triangulation of polygons, colour
packing 0xAABBGGRR, texture coordinates per triangle corner, skipped
elements and properties), round trips through write_ply, and the refusals.
"""
import struct

import numpy as np
import pytest

from j3d_amd import meshio

# a unit cube: 8 corners, 6 quads (outward, counter-clockwise)
CUBE_V = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
CUBE_F = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
# the fans (a, b, c), (a, c, d) of those quads, written out
CUBE_T = [(0, 3, 2), (0, 2, 1), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (1, 2, 6), (1, 6, 5),
          (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]


def _write(path, data):
    with open(path, "wb") as f:
        f.write(data)
    return str(path)


def test_ascii_quads_normals_colours(tmp_path):
    lines = ["ply", "format ascii 1.0", "comment a cube", "obj_info made by hand",
             "element vertex 8", "property float x", "property float y", "property float z",
             "property float nx", "property float ny", "property float nz",
             "property uchar red", "property uchar green", "property uchar blue", "property uchar alpha",
             "element face 6", "property list uchar int vertex_indices", "end_header"]
    for i, (x, y, z) in enumerate(CUBE_V):
        lines.append("%g %g %g %g %g %g %d %d %d %d" % (x, y, z, 0, 0, 1 if z else -1, 10 * i, 20 + i, 200, 255 - i))
    for f in CUBE_F:
        lines.append("4 " + " ".join(map(str, f)))
    m = meshio.read_ply(_write(tmp_path / "cube.ply", ("\n".join(lines) + "\n").encode()))
    assert np.array_equal(m["vertices"], np.array(CUBE_V, np.float32))
    assert np.array_equal(m["triangles"], np.array(CUBE_T, np.uint32))
    assert np.array_equal(m["normals"][:, 2], np.array([-1, -1, -1, -1, 1, 1, 1, 1], np.float32))
    # 0xAABBGGRR: red in the low byte
    want = [(10 * i) | (20 + i) << 8 | 200 << 16 | (255 - i) << 24 for i in range(8)]
    assert m["colors"].tolist() == want
    assert m["uv"] is None


def test_ascii_crlf_and_float_parsing(tmp_path):
    """Windows line ends; a float32 token is the float32 nearest its decimal value."""
    txt = ("ply\r\nformat ascii 1.0\r\nelement vertex 3\r\nproperty float x\r\nproperty float y\r\n"
           "property float z\r\nelement face 1\r\nproperty list uchar uint vertex_indices\r\nend_header\r\n"
           "0.1 1e-7 -3.4028234e38\r\n1 0 0\r\n0 1 0\r\n3 0 1 2\r\n")
    m = meshio.read_ply(_write(tmp_path / "crlf.ply", txt.encode()))
    assert m["vertices"][0].tolist() == [np.float32(0.1), np.float32(1e-7), np.float32(-3.4028234e38)]
    assert m["triangles"].tolist() == [[0, 1, 2]]


def test_binary_big_endian_doubles_skipped_parts(tmp_path):
    """Doubles for x y z, no alpha (255), an extra vertex property and an
    element between vertex and face that must be skipped, an int-counted
    uint list, a pentagon (3 triangles) and a two-corner face (dropped)."""
    hdr = ("ply\nformat binary_big_endian 1.0\nelement vertex 5\nproperty double x\nproperty double y\n"
           "property double z\nproperty float quality\nproperty uchar red\nproperty uchar green\n"
           "property uchar blue\nelement edge 2\nproperty int vertex1\nproperty int vertex2\n"
           "property list uchar short tag\nelement face 2\nproperty list int uint vertex_index\n"
           "property ushort flags\nend_header\n").encode()
    body = b""
    xyz = [(0.5, -1.25, 3.0), (1e-3, 2.0, -0.0), (7.0, 8.0, 9.0), (1.0 / 3.0, 0.0, 1.0), (-2.0, -4.0, -8.0)]
    for i, (x, y, z) in enumerate(xyz):
        body += struct.pack(">dddfBBB", x, y, z, 0.25 * i, 1 + i, 100 + i, 250 - i)
    body += struct.pack(">iiBhh", 0, 1, 2, -5, 6) + struct.pack(">iiB", 2, 3, 0)
    body += struct.pack(">i5IH", 5, 4, 3, 2, 1, 0, 7) + struct.pack(">i2IH", 2, 0, 1, 9)
    m = meshio.read_ply(_write(tmp_path / "be.ply", hdr + body))
    assert np.array_equal(m["vertices"], np.array(xyz, np.float64).astype(np.float32))
    assert m["triangles"].tolist() == [[4, 3, 2], [4, 2, 1], [4, 1, 0]]
    assert m["colors"].tolist() == [(1 + i) | (100 + i) << 8 | (250 - i) << 16 | 255 << 24 for i in range(5)]
    assert m["normals"] is None and m["uv"] is None


def test_binary_little_endian_face_texcoords(tmp_path):
    """texcoord lists (MeshLab's per-face layout): a triangle's 6 values as they
    are, a quad's 8 values split with its fan."""
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
           "property float z\nelement face 2\nproperty list uchar int vertex_indices\n"
           "property list uchar float texcoord\nend_header\n").encode()
    body = b"".join(struct.pack("<3f", *v) for v in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)])
    body += struct.pack("<B3iB6f", 3, 0, 1, 2, 6, 0.0, 0.0, 1.0, 0.0, 1.0, 1.0)
    body += struct.pack("<B4iB8f", 4, 0, 1, 2, 3, 8, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8)
    m = meshio.read_ply(_write(tmp_path / "tc.ply", hdr + body))
    assert m["triangles"].tolist() == [[0, 1, 2], [0, 1, 2], [0, 2, 3]]
    f32 = lambda *a: [float(np.float32(x)) for x in a]
    assert m["uv"].tolist() == [f32(0, 0, 1, 0, 1, 1), f32(0.1, 0.2, 0.3, 0.4, 0.5, 0.6),
                                f32(0.1, 0.2, 0.5, 0.6, 0.7, 0.8)]


def test_per_vertex_texture_coordinates(tmp_path):
    """s, t per vertex become each triangle corner's (u, v)."""
    txt = ("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
           "property float s\nproperty float t\nelement face 1\nproperty list uchar int vertex_indices\n"
           "end_header\n0 0 0 0 0\n1 0 0 0.5 0\n1 1 0 0.5 1\n0 1 0 0 1\n4 0 1 2 3\n")
    m = meshio.read_ply(_write(tmp_path / "st.ply", txt.encode()))
    assert m["uv"].tolist() == [[0, 0, 0.5, 0, 0.5, 1], [0, 0, 0.5, 1, 0, 1]]


@pytest.mark.parametrize("ascii", [False, True])
def test_round_trip(tmp_path, ascii):
    """write_ply then read_ply gives back every array bit for bit (ascii
    prints 9 significant digits, enough to round-trip a float32)."""
    import j3d_amd as J
    v, t = J.torus(60, 30, seed=4)
    rng = np.random.default_rng(1)
    n = rng.normal(size=v.shape).astype(np.float32)
    c = rng.integers(0, 2 ** 32, size=len(v), dtype=np.uint64).astype(np.uint32)
    uv = rng.uniform(-2, 2, size=(len(t), 6)).astype(np.float32)
    p = str(tmp_path / ("rt_a.ply" if ascii else "rt_b.ply"))
    meshio.write_ply(p, v, t, normals=n, colors=c, uv=uv, ascii=ascii)
    m = meshio.read_ply(p)
    assert np.array_equal(m["vertices"].view(np.uint32), v.view(np.uint32))
    assert np.array_equal(m["triangles"], t)
    assert np.array_equal(m["normals"].view(np.uint32), n.view(np.uint32))
    assert np.array_equal(m["colors"], c)
    assert np.array_equal(m["uv"].view(np.uint32), uv.view(np.uint32))
    # geometry only
    meshio.write_ply(p, v, t, ascii=ascii)
    m = meshio.read_ply(p)
    assert np.array_equal(m["vertices"], v) and np.array_equal(m["triangles"], t)
    assert m["normals"] is None and m["colors"] is None and m["uv"] is None


def test_binary_layout_of_write_ply(tmp_path):
    """The exact bytes write_ply produces for one triangle with colours."""
    p = str(tmp_path / "one.ply")
    meshio.write_ply(p, [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]], colors=[0x11223344, 0x55667788, 0xFF000000])
    data = open(p, "rb").read()
    hdr, body = data.split(b"end_header\n")
    assert hdr.startswith(b"ply\nformat binary_little_endian 1.0\n")
    assert b"element vertex 3\n" in hdr and b"element face 1\n" in hdr
    assert b"property uchar red\nproperty uchar green\nproperty uchar blue\nproperty uchar alpha\n" in hdr
    want = (struct.pack("<3f4B", 0, 0, 0, 0x44, 0x33, 0x22, 0x11) + struct.pack("<3f4B", 1, 0, 0, 0x88, 0x77, 0x66, 0x55)
            + struct.pack("<3f4B", 0, 1, 0, 0, 0, 0, 0xFF) + struct.pack("<B3i", 3, 0, 1, 2))
    assert body == want


def test_refusals(tmp_path):
    with pytest.raises(meshio.PlyError, match="cannot open"):
        meshio.read_ply(str(tmp_path / "missing.ply"))
    with pytest.raises(meshio.PlyError, match="magic"):
        meshio.read_ply(_write(tmp_path / "x.ply", b"plx\nformat ascii 1.0\nend_header\n"))
    with pytest.raises(meshio.PlyError, match="format"):
        meshio.read_ply(_write(tmp_path / "f.ply", b"ply\nformat binary_middle_endian 1.0\nend_header\n"))
    with pytest.raises(meshio.PlyError, match="end_header"):
        meshio.read_ply(_write(tmp_path / "h.ply", b"ply\nformat ascii 1.0\nelement vertex 1\n"))
    good = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
            "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n").encode()
    body = struct.pack("<9f", 0, 0, 0, 1, 0, 0, 0, 1, 0)
    meshio.read_ply(_write(tmp_path / "ok.ply", good + body + struct.pack("<B3i", 3, 0, 1, 2)))
    with pytest.raises(meshio.PlyError, match="truncated"):
        meshio.read_ply(_write(tmp_path / "t.ply", good + body + struct.pack("<B2i", 3, 0, 1)))
    bad_tok = b"ply\nformat ascii 1.0\nelement vertex 1\nproperty int x\nproperty int y\nproperty int z\nend_header\n1.5 0 0\n"
    with pytest.raises(meshio.PlyError, match="malformed vertex data"):
        meshio.read_ply(_write(tmp_path / "tok.ply", bad_tok))
    with pytest.raises(meshio.PlyError, match="refers to vertex 3"):
        meshio.read_ply(_write(tmp_path / "i.ply", good + body + struct.pack("<B3i", 3, 0, 1, 3)))
    with pytest.raises(meshio.PlyError, match="refers to vertex -1"):
        meshio.read_ply(_write(tmp_path / "n.ply", good + body + struct.pack("<B3i", 3, 0, 1, -1)))
    bad_tc = good.replace(b"end_header", b"property list uchar float texcoord\nend_header")
    with pytest.raises(meshio.PlyError, match="texcoord"):
        meshio.read_ply(_write(tmp_path / "c.ply", bad_tc + body + struct.pack("<B3iB4f", 3, 0, 1, 2, 4, 0, 0, 1, 1)))
    huge = good.replace(b"element vertex 3", b"element vertex 4000000000")
    with pytest.raises(meshio.PlyError, match="more instances than the file holds"):
        meshio.read_ply(_write(tmp_path / "v.ply", huge + body))
    many_faces = good.replace(b"element face 1", b"element face 1000")
    with pytest.raises(meshio.PlyError, match="than the file holds"):
        meshio.read_ply(_write(tmp_path / "m.ply", many_faces + body + struct.pack("<B3i", 3, 0, 1, 2)))
    with pytest.raises(meshio.PlyError, match="vertex 5 of 3"):
        meshio.write_ply(str(tmp_path / "w.ply"), [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 5]])
    with pytest.raises(ValueError):
        meshio.write_ply(str(tmp_path / "w.ply"), [[0, 0, 0]], [[0, 0, 0]], colors=[1, 2])
