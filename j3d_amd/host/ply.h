// ply.h -- read_ply / write_ply of j3d's io layer (j3d/io.h, j3d/io.cpp:733-741)
// for the canvas layer's flat mesh arrays.
//
// The reference forwards both calls to jtk::read_ply / jtk::write_ply, whose
// sources are not in the tree; this is the PLY format as published (Turk,
// "The PLY Polygon File Format"): a text header of elements and typed
// properties, then the data in ascii, binary_little_endian or
// binary_big_endian.
//
// read_ply fills
//   vertices   x y z of each vertex (3 floats per vertex; any scalar type in the file),
//   normals    nx ny nz per vertex when the file has them, else empty,
//   clrs       0xAABBGGRR per vertex from red/green/blue[/alpha] (or r/g/b/a;
//              integer channels as they are, float channels x 255; alpha 255
//              when absent), else empty,
//   triangles  3 vertex indices per triangle: the face element's
//              vertex_indices (or vertex_index) lists, a polygon of n > 3
//              corners as the fan (0, i, i+1); faces of fewer than 3 corners
//              are dropped,
//   uv         6 floats per triangle (u, v of its three corners) from a face
//              property list `texcoord` (6 floats per triangle face; for a
//              polygon, 2 per corner) or from per-vertex s/t, u/v or
//              texture_u/texture_v; else empty.
// Other elements and properties are skipped.  A face index outside the
// vertex range, a truncated file or a malformed header makes it return
// false; ply_error() then says why.
//
// write_ply writes binary_little_endian (ascii with `ascii`): vertex x y z
// as float, nx ny nz when normals are given, red green blue alpha (uchar)
// when colours are given; one face list (uchar count, int indices) per
// triangle, and a `texcoord` list (uchar count, 6 floats) when uv is given.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace j3d_amd {

bool read_ply(const char* filename, std::vector<float>& vertices, std::vector<float>& normals,
              std::vector<uint32_t>& clrs, std::vector<uint32_t>& triangles, std::vector<float>& uv);
bool write_ply(const char* filename, const std::vector<float>& vertices, const std::vector<float>& normals,
               const std::vector<uint32_t>& clrs, const std::vector<uint32_t>& triangles,
               const std::vector<float>& uv, bool ascii = false);
// why the last read_ply / write_ply of this thread failed
const std::string& ply_error();

}  // namespace j3d_amd
