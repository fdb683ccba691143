// bvh_build.h -- host BVH builder and GPU re-pack (replaces the jtk::qbvh
// build at j3d/scene.cpp:23 with this library's own builder).
//
// Device layout (all arrays 16-byte aligned, one node per cache line):
//
//   wide node, W = 4 (128 B) or 8 (256 B), float words:
//     [0W..1W) lo.x  [1W..2W) hi.x  [2W..3W) lo.y  [3W..4W) hi.y
//     [4W..5W) lo.z  [5W..6W) hi.z  [6W..7W) child refs (uint32)  [7W..8W) unused
//   child ref: internal child  -> node index (bit 31 clear)
//              leaf            -> 0x80000000 | index of the leaf's first triangle record
//              empty slot      -> 0xFFFFFFFF with box lo = +inf, hi = -inf (never entered)
//   Child boxes are widened by 2^-20 of their coordinate magnitude so that the
//   traversal's box culling is conservative.
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
  int leaf_size = 4;  // max triangles per leaf
  int threads = 0;    // 0 = hardware_concurrency
};

struct BuiltBVH {
  int width = 4;
  std::vector<float> nodes;  // n_nodes * 8 * width floats
  uint32_t n_nodes = 0;
  std::vector<float> tris;   // n_tris * 12 floats
  uint32_t max_depth = 0;    // wide levels, root = 1
  uint32_t max_stack = 0;    // upper bound on traversal stack entries
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of all triangles (unwidened)
};

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0xFFFFFFFFu;

// Builds and re-packs.  Returns false on invalid input (index out of range,
// unsupported width, more than 2^31-2 triangles).
bool build_bvh(const float* xyz, uint64_t n_vertices, const uint32_t* tris, uint64_t n_tris, const BuildOptions& opt,
               BuiltBVH& out, const char** err);

}  // namespace j3d
