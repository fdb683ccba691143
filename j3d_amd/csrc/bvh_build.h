// bvh_build.h -- host BVH builder and GPU re-pack (replaces the jtk::qbvh
// build at j3d/scene.cpp:23 with this library's own builder).
//
// Device layout (16-byte aligned arrays):
//
//   quantized wide node, W = 4 (64 B, four 16-byte loads) or 8 (96 B, six):
//     word 0..2  p = node origin (float x, y, z): the minimum of the children's boxes
//     word 3     biased exponents e_x | e_y << 8 | e_z << 16: the quantum of axis k is 2^(e_k - 127)
//     W = 4:  word 4 qlo.x[4]  5 qhi.x[4]  6 qlo.y  7 qhi.y  8 qlo.z  9 qhi.z   (one byte per child)
//             word 10..13 child refs, 14..15 unused
//     W = 8:  words 4..7 qlo.x[8], qhi.x[8]; 8..11 y; 12..15 z;  words 16..23 child refs
//   child box on axis k = [p_k + qlo * 2^e, p_k + qhi * 2^e]; qlo is rounded
//   down and qhi up from the child's box widened by 2^-20 of its magnitude,
//   so the traversal's box culling is conservative.  Empty slots have
//   qlo = 255, qhi = 0 (never entered) and ref 0xFFFFFFFF.
//   child ref: internal child  -> node index (bit 31 clear)
//              leaf            -> 0x80000000 | (count - 1) << 29 | index of the leaf's first
//                                 triangle record (count 1..4, index < 2^29)
//
//   triangle record (48 B), in leaf order:
//     float4(v0.xyz, id)   id = original triangle index (uint32 bits)
//     float4(e1.xyz, flag) e1 = v1 - v0; flag bit 0 = last triangle of its leaf
//     float4(e2.xyz, 0)    e2 = v2 - v0
#pragma once
#include <stdint.h>

#include <vector>

namespace j3d {

struct BuildOptions {
  int width = 4;      // 4 or 8
  int leaf_size = 4;  // max triangles per leaf (1..4)
  int threads = 0;    // 0 = hardware_concurrency
};

// 32-bit words per node.
constexpr int node_words(int width) { return width == 8 ? 24 : 16; }

struct BuiltBVH {
  int width = 4;
  std::vector<uint32_t> nodes;  // n_nodes * node_words(width)
  uint32_t n_nodes = 0;
  std::vector<float> tris;   // n_tris * 12 floats
  uint32_t max_depth = 0;    // wide levels, root = 1
  uint32_t max_stack = 0;    // upper bound on traversal stack entries
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of all triangles (unwidened)
};

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafIndexMask = (1u << 29) - 1u;
constexpr int kLeafCountShift = 29;
constexpr uint32_t kEmptyRef = 0xFFFFFFFFu;

// Builds and re-packs.  Returns false on invalid input (index out of range,
// unsupported width, more than 2^31-2 triangles).
bool build_bvh(const float* xyz, uint64_t n_vertices, const uint32_t* tris, uint64_t n_tris, const BuildOptions& opt,
               BuiltBVH& out, const char** err);

}  // namespace j3d
