// bvh_gpu.hip -- device build of the traversal BVH (replaces the host build of
// jtk::qbvh at j3d/scene.cpp:23 when a caller asks for it), plus the per-mesh
// data the host upload also derives: triangle normals (scene.cpp:20) and the
// hit completion records.
//
// Linear BVH: triangles sorted by the 63-bit Morton code of their box centre
// in the mesh's bounds; a binary radix tree over the sorted codes (each
// internal node found independently from the longest common prefixes of its
// neighbours); node boxes bottom-up, one launch per tree level (the level
// lists come from a top-down pass; kernel boundaries order the levels, so no
// thread waits for or fences against another); the levels above a frontier of
// some 16,000-32,000 subtrees rebuilt with a full-sweep SAH on the host over
// those subtrees' boxes; then, level by level from the root, each binary node
// is opened into up to W children (the largest-area internal child first, as
// the host collapser does), and written as the same quantized wide node, leaf
// encoding and 48-byte triangle records that build_bvh produces.  A subtree
// of at most leaf_size triangles becomes a leaf: its triangles are contiguous
// in Morton order.  The tree is of lower quality than the host's binned SAH
// (more nodes visited per ray) and builds far faster; results do not depend
// on it (closest hit is the minimum of (t, object, triangle)).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "bvh_build.h"
#include "j3dg.h"
#include "j3dg_internal.h"
#include "rt_math.h"

namespace j3d {
namespace lbvh {

// Binary radix tree node.  Child refs: bit 31 set = a single triangle (its
// sorted position), else an internal node index.
struct BNode {
  uint32_t left, right;
  uint32_t first, last;  // sorted positions covered (inclusive)
  float lo[3], hi[3];
};
constexpr uint32_t kPrimBit = 0x80000000u;

// Work item of the level-by-level collapse: binary node -> wide node slot.
struct Item {
  uint32_t bnode, wslot;
  uint32_t stack;  // traversal stack entries the ancestors may have pushed
};

// The 21 low bits of v moved to every third bit.
__device__ inline uint64_t spread21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | v << 32) & 0x1F00000000FFFFull;
  v = (v | v << 16) & 0x1F0000FF0000FFull;
  v = (v | v << 8) & 0x100F00F00F00F00Full;
  v = (v | v << 4) & 0x10C30C30C30C30C3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// Per triangle: its box (6 floats, lo then hi) and the Morton code of the box
// centre; the sort values start as the triangle indices.
__global__ void k_prims(const float* __restrict__ xyz, const uint32_t* __restrict__ tris, uint32_t n, float3 base,
                        float3 inv_ext, float* __restrict__ box, uint64_t* __restrict__ code,
                        uint32_t* __restrict__ index) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* a = xyz + 3 * (size_t)tris[3 * (size_t)i];
  const float* b = xyz + 3 * (size_t)tris[3 * (size_t)i + 1];
  const float* c = xyz + 3 * (size_t)tris[3 * (size_t)i + 2];
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = fminf(a[k], fminf(b[k], c[k]));
    hi[k] = fmaxf(a[k], fmaxf(b[k], c[k]));
  }
  float* bx = box + 6 * (size_t)i;
  for (int k = 0; k < 3; ++k) bx[k] = lo[k], bx[3 + k] = hi[k];
  const float cx = ((lo[0] + hi[0]) * 0.5f - base.x) * inv_ext.x;
  const float cy = ((lo[1] + hi[1]) * 0.5f - base.y) * inv_ext.y;
  const float cz = ((lo[2] + hi[2]) * 0.5f - base.z) * inv_ext.z;
  auto q = [](float t) {
    t = t * 2097152.f;
    return t < 0.f ? 0ull : (t > 2097151.f ? 2097151ull : (uint64_t)t);
  };
  code[i] = spread21(q(cx)) << 2 | spread21(q(cy)) << 1 | spread21(q(cz));
  index[i] = i;
}

// Length of the common prefix of sorted keys i and j (ties broken by
// position), -1 outside [0, n).
__device__ inline int delta(const uint64_t* __restrict__ code, uint32_t n, int i, int j) {
  if (j < 0 || j >= (int)n) return -1;
  const uint64_t a = code[i], b = code[j];
  if (a != b) return __clzll(a ^ b);
  return 64 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Internal node i of the radix tree over n >= 2 sorted keys.
__global__ void k_radix_tree(const uint64_t* __restrict__ code, uint32_t n, BNode* __restrict__ nodes) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= (int)n - 1) return;
  const int d = delta(code, n, i, i + 1) - delta(code, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = delta(code, n, i, i - d);
  int lmax = 2;
  while (delta(code, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(code, n, i, j);
  int s = 0;
  for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
    if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  BNode& nd = nodes[i];
  nd.first = (uint32_t)first;
  nd.last = (uint32_t)last;
  nd.left = first == gamma ? kPrimBit | (uint32_t)gamma : (uint32_t)gamma;
  nd.right = last == gamma + 1 ? kPrimBit | (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
}

__device__ inline void child_box(const BNode* nodes, const float* box, const uint32_t* sorted, uint32_t ref,
                                 float* lo, float* hi) {
  if (ref & kPrimBit) {
    const float* b = box + 6 * (size_t)sorted[ref & ~kPrimBit];
    for (int k = 0; k < 3; ++k) lo[k] = b[k], hi[k] = b[3 + k];
  } else {
    const BNode& c = nodes[ref];
    for (int k = 0; k < 3; ++k) lo[k] = c.lo[k], hi[k] = c.hi[k];
  }
}

// Top-down level lists: the internal children of one level's nodes, appended
// to the next level's list.
__global__ void k_level_down(const BNode* __restrict__ nodes, const uint32_t* __restrict__ in, uint32_t n_in,
                             uint32_t* __restrict__ out, uint32_t cap, uint32_t* __restrict__ count) {
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= n_in) return;
  const BNode& nd = nodes[in[it]];
  const uint32_t l = nd.left, r = nd.right;
  const uint32_t k = (l & kPrimBit) ? 0u : 1u;
  const uint32_t m = k + ((r & kPrimBit) ? 0u : 1u);
  if (!m) return;
  const uint32_t at = atomicAdd(count, m);
  if (at + m > cap) return;  // only a malformed tree; the host sees the count and fails
  if (k) out[at] = l;
  if (!(r & kPrimBit)) out[at + k] = r;
}

// Boxes of one level's nodes from their children (triangles, or nodes of the
// level below, written by the previous launch).
__global__ void k_level_boxes(BNode* __restrict__ nodes, const float* __restrict__ box,
                              const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ list, uint32_t n_in) {
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= n_in) return;
  BNode& nd = nodes[list[it]];
  float l0[3], h0[3], l1[3], h1[3];
  child_box(nodes, box, sorted, nd.left, l0, h0);
  child_box(nodes, box, sorted, nd.right, l1, h1);
  for (int k = 0; k < 3; ++k) {
    nd.lo[k] = fminf(l0[k], l1[k]);
    nd.hi[k] = fmaxf(h0[k], h1[k]);
  }
}

// The host builder's child box widening (2^-20 of the magnitude).
__device__ inline void widen(const float* blo, const float* bhi, float* lo, float* hi) {
  for (int k = 0; k < 3; ++k) {
    const float m = fmaxf(fabsf(blo[k]), fabsf(bhi[k]));
    const float e = m * (1.0f / 1048576.0f) + 1e-30f;
    lo[k] = blo[k] - e;
    hi[k] = bhi[k] + e;
  }
}

// Quantized wide node, the layout of bvh_build.h: the decoded child boxes
// contain the widened boxes (floor / ceil, checked in double).
template <int W>
__device__ void write_wide_node(uint32_t* nd, int n, const float (*lo)[3], const float (*hi)[3],
                                const uint32_t* refs) {
  constexpr int kWords = W == 8 ? 24 : 16;
  for (int i = 0; i < kWords; ++i) nd[i] = 0;
  float p[3], nh[3];
  for (int k = 0; k < 3; ++k) {
    p[k] = INFINITY;
    nh[k] = -INFINITY;
    for (int i = 0; i < n; ++i) {
      p[k] = fminf(p[k], lo[i][k]);
      nh[k] = fmaxf(nh[k], hi[i][k]);
    }
    if (n == 0) p[k] = nh[k] = 0.f;
  }
  uint32_t eb[3];
  uint32_t qlo[3][8], qhi[3][8];
  for (int k = 0; k < 3; ++k) {
    const double ext = (double)nh[k] - (double)p[k];
    int e = -126;
    if (ext > 0.0) {
      e = (int)ceil(log2(ext / 255.0));
      if (e < -126) e = -126;
      while ((double)p[k] + 255.0 * ldexp(1.0, e) < (double)nh[k]) ++e;
    }
    eb[k] = (uint32_t)(e + 127);
    const double sc = ldexp(1.0, e);
    for (int i = 0; i < W; ++i) {
      if (i >= n) {
        qlo[k][i] = 255;
        qhi[k][i] = 0;
        continue;
      }
      double ql = floor(((double)lo[i][k] - (double)p[k]) / sc);
      double qh = ceil(((double)hi[i][k] - (double)p[k]) / sc);
      while (ql > 0 && (double)p[k] + ql * sc > (double)lo[i][k]) ql -= 1;
      while ((double)p[k] + qh * sc < (double)hi[i][k]) qh += 1;
      qlo[k][i] = (uint32_t)fmin(fmax(ql, 0.0), 255.0);
      qhi[k][i] = (uint32_t)fmin(fmax(qh, 0.0), 255.0);
    }
  }
  nd[0] = __float_as_uint(p[0]);
  nd[1] = __float_as_uint(p[1]);
  nd[2] = __float_as_uint(p[2]);
  nd[3] = eb[0] | (eb[1] << 8) | (eb[2] << 16);
  auto pack4 = [&](const uint32_t (*q)[8], int k, int f) {
    return q[k][f] | (q[k][f + 1] << 8) | (q[k][f + 2] << 16) | (q[k][f + 3] << 24);
  };
  if (W == 4) {
    nd[4] = pack4(qlo, 0, 0), nd[5] = pack4(qhi, 0, 0);
    nd[6] = pack4(qlo, 1, 0), nd[7] = pack4(qhi, 1, 0);
    nd[8] = pack4(qlo, 2, 0), nd[9] = pack4(qhi, 2, 0);
    for (int i = 0; i < 4; ++i) nd[10 + i] = i < n ? refs[i] : kEmptyRef;
  } else {
    for (int k = 0; k < 3; ++k) {
      nd[4 + 4 * k] = pack4(qlo, k, 0), nd[5 + 4 * k] = pack4(qlo, k, 4);
      nd[6 + 4 * k] = pack4(qhi, k, 0), nd[7 + 4 * k] = pack4(qhi, k, 4);
    }
    for (int i = 0; i < 8; ++i) nd[16 + i] = i < n ? refs[i] : kEmptyRef;
  }
}

struct CollapseArgs {
  const BNode* nodes;
  const float* box;
  const uint32_t* sorted;
  uint32_t leaf_size;
  const Item* in;
  uint32_t n_in;
  Item* out;
  uint32_t* counters;  // [0] items written to `out`, [1] wide nodes allocated, [2] overflow flag,
                       // [3] largest stack bound seen
  uint32_t wide_cap;
  uint32_t* wide;      // wide nodes, kWords words each
  uint8_t* last;       // per sorted position: last triangle of its leaf
};

__device__ inline uint32_t ref_count(const BNode* nodes, uint32_t ref) {
  return (ref & kPrimBit) ? 1u : nodes[ref].last - nodes[ref].first + 1u;
}
__device__ inline uint32_t ref_first(const BNode* nodes, uint32_t ref) {
  return (ref & kPrimBit) ? (ref & ~kPrimBit) : nodes[ref].first;
}

// One level of the collapse: every item opens its binary node into up to W
// children and writes its wide node; internal children go to the next level.
template <int W>
__global__ void k_collapse(CollapseArgs a) {
  constexpr int kWords = W == 8 ? 24 : 16;
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= a.n_in) return;
  const Item item = a.in[it];
  uint32_t ch[W];
  int n = 1;
  ch[0] = item.bnode;
  // open the largest-area child that is an internal node of more than
  // leaf_size triangles, until W children or none is left to open
  for (;;) {
    int best = -1;
    float best_area = -1.f;
    for (int i = 0; i < n; ++i) {
      if ((ch[i] & kPrimBit) || ref_count(a.nodes, ch[i]) <= a.leaf_size) continue;
      const BNode& c = a.nodes[ch[i]];
      const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
      float area = dx * dy + dy * dz + dz * dx;
      if (!(area >= 0.f)) area = 0.f;  // a NaN box (NaN vertices) is still opened, so every level descends
      if (area > best_area) best_area = area, best = i;
    }
    if (best < 0) break;
    const uint32_t c = ch[best];
    ch[best] = a.nodes[c].left;
    ch[n++] = a.nodes[c].right;
    if (n == W) break;
  }
  // the host collapser's bound: entries pushed on the way down plus n - 1 here
  const uint32_t pushed = item.stack + (uint32_t)(n - 1);
  atomicMax(&a.counters[3], pushed);
  uint32_t n_int = 0;
  for (int i = 0; i < n; ++i)
    if (!(ch[i] & kPrimBit) && ref_count(a.nodes, ch[i]) > a.leaf_size) ++n_int;
  uint32_t base = 0;
  if (n_int) {
    base = atomicAdd(&a.counters[1], n_int);
    if (base + n_int > a.wide_cap) {
      atomicOr(&a.counters[2], 1u);
      return;
    }
  }
  const uint32_t obase = n_int ? atomicAdd(&a.counters[0], n_int) : 0u;
  float lo[W][3], hi[W][3];
  uint32_t refs[W];
  uint32_t k = 0;
  for (int i = 0; i < n; ++i) {
    float bl[3], bh[3];
    child_box(a.nodes, a.box, a.sorted, ch[i], bl, bh);
    widen(bl, bh, lo[i], hi[i]);
    const uint32_t cnt = ref_count(a.nodes, ch[i]);
    if (cnt <= a.leaf_size) {
      const uint32_t first = ref_first(a.nodes, ch[i]);
      refs[i] = kLeafBit | ((cnt - 1u) << kLeafCountShift) | first;
      a.last[first + cnt - 1u] = 1;
    } else {
      refs[i] = base + k;
      a.out[obase + k] = Item{ch[i], base + k, pushed};
      ++k;
    }
  }
  write_wide_node<W>(a.wide + (size_t)item.wslot * kWords, n, lo, hi, refs);
}

// Triangle records in Morton (leaf) order: float4(v0, id), float4(e1, last),
// float4(e2, 0) -- the operands build_bvh writes.
__global__ void k_tri_records(const float* __restrict__ xyz, const uint32_t* __restrict__ tris,
                              const uint32_t* __restrict__ sorted, const uint8_t* __restrict__ last, uint32_t n,
                              float4* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t p = sorted[s];
  const float* v0 = xyz + 3 * (size_t)tris[3 * (size_t)p];
  const float* v1 = xyz + 3 * (size_t)tris[3 * (size_t)p + 1];
  const float* v2 = xyz + 3 * (size_t)tris[3 * (size_t)p + 2];
  float4* r = out + 3 * (size_t)s;
  r[0] = make_float4(v0[0], v0[1], v0[2], __uint_as_float(p));
  r[1] = make_float4(v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2], __uint_as_float((uint32_t)last[s]));
  r[2] = make_float4(v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2], 0.f);
}

// Per-triangle normals (the host upload's triangle_normal, same rounding) and
// the hit completion records (v0, n.x), (v1, n.y), (v2, n.z).
__global__ void k_normals_fin(const float* __restrict__ xyz, const uint32_t* __restrict__ tris, uint32_t n,
                              float* __restrict__ normals, float4* __restrict__ fin) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* v[3];
  for (int k = 0; k < 3; ++k) v[k] = xyz + 3 * (size_t)tris[3 * (size_t)i + k];
  float nn[3];
  triangle_normal(v[0], v[1], v[2], nn);
  for (int k = 0; k < 3; ++k) normals[3 * (size_t)i + k] = nn[k];
  for (int k = 0; k < 3; ++k) fin[3 * (size_t)i + k] = make_float4(v[k][0], v[k][1], v[k][2], nn[k]);
}

// ---- SAH over the top levels ------------------------------------------
//
// Morton order splits the top of the tree at spatial medians, where a bad
// split costs the most.  The tree is cut at the frontier of maximal subtrees
// of at most T = n / 2^14 triangles (some 16,000-32,000 clusters); the levels above it are
// rebuilt on the host with a full-sweep SAH over the clusters' boxes, weighted
// by their triangle counts, and written back into the binary nodes they
// replace (exactly as many: a binary tree over m clusters has m - 1 internal
// nodes), the root staying node 0.  A rebuilt node's triangles are not
// contiguous in Morton order, so it is given a count above leaf_size: the
// collapse always opens it.

// Internal nodes of more than T triangles (the part rebuilt) and their
// children of at most T (the frontier).
__global__ void k_frontier(const BNode* __restrict__ nodes, uint32_t n, uint32_t T, uint32_t* __restrict__ frontier,
                           uint32_t* __restrict__ top, uint32_t* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  if (nodes[i].last - nodes[i].first + 1u <= T) return;
  top[atomicAdd(&counters[1], 1u)] = i;
  const uint32_t ch[2] = {nodes[i].left, nodes[i].right};
  for (int k = 0; k < 2; ++k)
    if (ref_count(nodes, ch[k]) <= T) frontier[atomicAdd(&counters[0], 1u)] = ch[k];
}

// Box and triangle count of each frontier entry: 8 words (lo, hi, count, ref).
__global__ void k_frontier_boxes(const BNode* __restrict__ nodes, const float* __restrict__ box,
                                 const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ frontier,
                                 uint32_t m, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float lo[3], hi[3];
  child_box(nodes, box, sorted, frontier[i], lo, hi);
  float* o = out + 8 * (size_t)i;
  for (int k = 0; k < 3; ++k) o[k] = lo[k], o[3 + k] = hi[k];
  o[6] = __uint_as_float(ref_count(nodes, frontier[i]));
  o[7] = __uint_as_float(frontier[i]);
}

// Writes rebuilt nodes: nodes[index[j]] = val[j].
__global__ void k_put_nodes(BNode* __restrict__ nodes, const uint32_t* __restrict__ index,
                            const BNode* __restrict__ val, uint32_t k) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) nodes[index[j]] = val[j];
}

constexpr bool kSahTop = true;

struct Cluster {
  float lo[3], hi[3];
  float c[3];
  uint32_t count, ref;
};

inline float half_area(const float* lo, const float* hi) {
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  const float a = dx * dy + dy * dz + dz * dx;
  return a >= 0.f ? a : 0.f;  // NaN boxes cost nothing (they hold no hits)
}

// Full-sweep SAH over cl[a, b) (b - a >= 2); appends the subtree's internal
// nodes to `out` (its root first) and returns the root's position there.
// Children refs: a cluster's own ref, or kPrimBit-free positions in `out`
// tagged by `tag` (resolved to node indices by the caller).
struct SahNode {
  uint32_t left, right;  // cluster ref, or (kSahTag | position in out)
  uint32_t count;
  float lo[3], hi[3];
};
constexpr uint32_t kSahTag = 0x40000000u;

inline uint32_t sah_build(std::vector<Cluster>& cl, uint32_t a, uint32_t b, std::vector<SahNode>& out,
                          std::vector<float>& right_area) {
  const uint32_t pos = (uint32_t)out.size();
  out.emplace_back();
  SahNode nd;
  nd.count = 0;
  for (int k = 0; k < 3; ++k) nd.lo[k] = INFINITY, nd.hi[k] = -INFINITY;
  for (uint32_t i = a; i < b; ++i) {
    nd.count += cl[i].count;
    for (int k = 0; k < 3; ++k) nd.lo[k] = fminf(nd.lo[k], cl[i].lo[k]), nd.hi[k] = fmaxf(nd.hi[k], cl[i].hi[k]);
  }
  // best split over the three centroid orders
  int best_axis = -1;
  uint32_t best_mid = a + (b - a) / 2;
  double best_cost = INFINITY;
  for (int axis = 0; axis < 3; ++axis) {
    std::sort(cl.begin() + a, cl.begin() + b, [axis](const Cluster& x, const Cluster& y) {
      return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.ref < y.ref);
    });
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double cnt = 0.0;
    for (uint32_t i = b; i-- > a + 1;) {  // right_area[i]: clusters [i, b)
      for (int k = 0; k < 3; ++k) lo[k] = fminf(lo[k], cl[i].lo[k]), hi[k] = fmaxf(hi[k], cl[i].hi[k]);
      cnt += cl[i].count;
      right_area[i] = (float)((double)half_area(lo, hi) * cnt);
    }
    for (int k = 0; k < 3; ++k) lo[k] = INFINITY, hi[k] = -INFINITY;
    cnt = 0.0;
    for (uint32_t i = a; i + 1 < b; ++i) {  // split before i + 1
      for (int k = 0; k < 3; ++k) lo[k] = fminf(lo[k], cl[i].lo[k]), hi[k] = fmaxf(hi[k], cl[i].hi[k]);
      cnt += cl[i].count;
      const double cost = (double)half_area(lo, hi) * cnt + (double)right_area[i + 1];
      // ties go to the split nearest the middle (equal boxes: a balanced tree)
      const uint32_t centre = a + (b - a) / 2;
      const uint32_t dm = i + 1 > centre ? i + 1 - centre : centre - (i + 1);
      const uint32_t db = best_mid > centre ? best_mid - centre : centre - best_mid;
      if (cost < best_cost || (cost == best_cost && dm < db)) best_cost = cost, best_axis = axis, best_mid = i + 1;
    }
  }
  if (best_axis != 2)  // the last sort was by z
    std::sort(cl.begin() + a, cl.begin() + b, [best_axis](const Cluster& x, const Cluster& y) {
      const int ax = best_axis < 0 ? 0 : best_axis;
      return x.c[ax] < y.c[ax] || (x.c[ax] == y.c[ax] && x.ref < y.ref);
    });
  auto side = [&](uint32_t lo_i, uint32_t hi_i) {
    return hi_i - lo_i == 1 ? cl[lo_i].ref : (kSahTag | sah_build(cl, lo_i, hi_i, out, right_area));
  };
  nd.left = side(a, best_mid);
  nd.right = side(best_mid, b);
  out[pos] = nd;
  return pos;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace lbvh

// Device upload and build (j3dg_upload_mesh with J3DG_BUILD_DEVICE).  Fills
// m's device arrays and info; the vertex and index arrays are copied first.
int device_build_mesh(Ctx* ctx, Mesh& m, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                      uint64_t n_triangles, int width, int leaf_size, double* upload_s) {
  using namespace lbvh;
  if (n_triangles < 2 || n_triangles > kLeafIndexMask)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "the device build needs 2 .. 2^29-1 triangles");
  if (width != 4 && width != 8) return ctx->fail(J3DG_ERR_INVALID_ARG, "BVH width must be 4 or 8");
  if (leaf_size < 1 || leaf_size > 4) return ctx->fail(J3DG_ERR_INVALID_ARG, "leaf size must be 1 .. 4");
  hipStream_t st = ctx->stream;
  const uint32_t n = (uint32_t)n_triangles;
  // indices in range (the host build checks the same)
  {
    const int threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<int> bad(threads, 0);
    std::vector<std::thread> pool;
    const uint64_t total = 3 * n_triangles, chunk = (total + threads - 1) / threads;
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&, t]() {
        const uint64_t a = std::min(total, t * chunk), b = std::min(total, (t + 1) * chunk);
        for (uint64_t i = a; i < b; ++i)
          if (tris[i] >= n_vertices) bad[t] = 1;
      });
    for (auto& th : pool) th.join();
    for (int b : bad)
      if (b) return ctx->fail(J3DG_ERR_INVALID_ARG, "triangle index out of range");
  }
  const int words = node_words(width);
  struct Tmp {
    void* p = nullptr;
    ~Tmp() {
      if (p) hipFree(p);
    }
  } box, code, code2, index, index2, bnodes, order, q0, q1, counters, sortTmp, last, wide;
  auto alloc = [&](Tmp& t, size_t bytes) { return hipMalloc(&t.p, std::max<size_t>(bytes, 16)) == hipSuccess; };
  const size_t fb = 3 * (size_t)n * sizeof(float4);
  const size_t tb = 3 * (size_t)n * sizeof(float4);
  bool ok = hipMalloc(&m.d_verts, std::max<size_t>(3 * n_vertices * sizeof(float), 16)) == hipSuccess &&
            hipMalloc(&m.d_idx, 3 * (size_t)n * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&m.d_normals, 3 * (size_t)n * sizeof(float)) == hipSuccess &&
            hipMalloc(&m.d_fin, fb + 3 * sizeof(float4)) == hipSuccess &&
            hipMalloc(&m.d_tris, tb + 3 * sizeof(float4)) == hipSuccess && alloc(box, 6 * (size_t)n * 4) &&
            alloc(code, (size_t)n * 8) && alloc(code2, (size_t)n * 8) && alloc(index, (size_t)n * 4) &&
            alloc(index2, (size_t)n * 4) && alloc(bnodes, (size_t)(n - 1) * sizeof(BNode)) &&
            alloc(order, (size_t)(n - 1) * 4) &&
            alloc(q0, (size_t)n * sizeof(Item)) && alloc(q1, (size_t)n * sizeof(Item)) && alloc(counters, 64) &&
            alloc(last, ((size_t)n + 3) / 4 * 4) && alloc(wide, (size_t)n * words * 4);
  if (!ok) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device allocation for the device BVH build failed");
  const auto tc = std::chrono::steady_clock::now();
  HIP_TRY(hipMemcpyAsync(m.d_verts, xyz, 3 * n_vertices * sizeof(float), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
  HIP_TRY(hipMemcpyAsync(m.d_idx, tris, 3 * (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
  HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
  *upload_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
  HIP_TRY(hipMemsetAsync(reinterpret_cast<char*>(m.d_fin) + fb, 0, 3 * sizeof(float4), st), J3DG_ERR_LAUNCH);
  HIP_TRY(hipMemsetAsync(reinterpret_cast<char*>(m.d_tris) + tb, 0, 3 * sizeof(float4), st), J3DG_ERR_LAUNCH);
  const uint32_t gb = (n + 255) / 256;
  hipLaunchKernelGGL(k_normals_fin, dim3(gb), dim3(256), 0, st, (const float*)m.d_verts, (const uint32_t*)m.d_idx, n,
                     m.d_normals, m.d_fin);
  // Morton codes in the mesh's bounds
  const float* lo = m.info.bb_min;
  const float* hi = m.info.bb_max;
  float3 base = make_float3(lo[0], lo[1], lo[2]);
  float3 inv = make_float3(hi[0] > lo[0] ? 1.f / (hi[0] - lo[0]) : 0.f, hi[1] > lo[1] ? 1.f / (hi[1] - lo[1]) : 0.f,
                           hi[2] > lo[2] ? 1.f / (hi[2] - lo[2]) : 0.f);
  hipLaunchKernelGGL(k_prims, dim3(gb), dim3(256), 0, st, (const float*)m.d_verts, (const uint32_t*)m.d_idx, n, base,
                     inv, (float*)box.p, (uint64_t*)code.p, (uint32_t*)index.p);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  // sort (code, index) pairs
  hipcub::DoubleBuffer<uint64_t> keys((uint64_t*)code.p, (uint64_t*)code2.p);
  hipcub::DoubleBuffer<uint32_t> vals((uint32_t*)index.p, (uint32_t*)index2.p);
  size_t sort_bytes = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, vals, (int)n, 0, 63, st), J3DG_ERR_LAUNCH);
  if (!alloc(sortTmp, sort_bytes)) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device sort storage");
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(sortTmp.p, sort_bytes, keys, vals, (int)n, 0, 63, st), J3DG_ERR_LAUNCH);
  const uint64_t* skeys = keys.Current();
  const uint32_t* sorted = vals.Current();
  // radix tree and boxes
  BNode* bn = (BNode*)bnodes.p;
  hipLaunchKernelGGL(k_radix_tree, dim3((n - 1 + 255) / 256), dim3(256), 0, st, skeys, n, bn);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  // level lists, root first: level L is order[level_at[L] .. level_at[L + 1])
  uint32_t* ord = (uint32_t*)order.p;
  uint32_t* cnt = (uint32_t*)counters.p;
  std::vector<uint32_t> level_at = {0u, 1u};
  {
    const uint32_t root = 0u;
    HIP_TRY(hipMemcpyAsync(ord, &root, sizeof(root), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
    for (;;) {
      const uint32_t a = level_at[level_at.size() - 2], b = level_at.back();
      if (a == b) break;
      if (b > n - 1 || level_at.size() > 100)
        return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: malformed radix tree");
      HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st), J3DG_ERR_LAUNCH);
      hipLaunchKernelGGL(k_level_down, dim3((b - a + 255) / 256), dim3(256), 0, st, (const BNode*)bn, ord + a, b - a,
                         ord + b, n - 1 - b, cnt);
      HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
      uint32_t c = 0;
      HIP_TRY(hipMemcpyAsync(&c, cnt, sizeof(c), hipMemcpyDeviceToHost, st), J3DG_ERR_LAUNCH);
      HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
      if (c > n - 1 - b) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: malformed radix tree");
      level_at.push_back(b + c);
    }
    if (level_at.back() != n - 1) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: malformed radix tree");
  }
  // boxes, deepest level first
  for (size_t L = level_at.size() - 1; L-- > 0;) {
    const uint32_t a = level_at[L], b = level_at[L + 1];
    if (a == b) continue;
    hipLaunchKernelGGL(k_level_boxes, dim3((b - a + 255) / 256), dim3(256), 0, st, bn, (const float*)box.p, sorted,
                       (const uint32_t*)(ord + a), b - a);
  }
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>((n + 255) / 256, 4096)), dim3(256), 0, st,
                     reinterpret_cast<uint32_t*>(last.p), (uint64_t)(n + 3) / 4, 0u);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  // SAH over the levels above a frontier of 16k-32k clusters (meshes of 8192 triangles or more)
  if (kSahTop && n >= 8192) {
    const uint32_t T = std::max<uint32_t>(n >> 14, 2u);
    uint32_t* frontier = reinterpret_cast<uint32_t*>(q0.p);  // n entries
    uint32_t* top = reinterpret_cast<uint32_t*>(order.p);    // n - 1 entries
    uint32_t* cnt2 = reinterpret_cast<uint32_t*>(counters.p);
    HIP_TRY(hipMemsetAsync(cnt2, 0, 2 * sizeof(uint32_t), st), J3DG_ERR_LAUNCH);
    hipLaunchKernelGGL(k_frontier, dim3((n - 1 + 255) / 256), dim3(256), 0, st, (const BNode*)bn, n, T, frontier, top,
                       cnt2);
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    uint32_t mk[2] = {0u, 0u};
    HIP_TRY(hipMemcpyAsync(mk, cnt2, sizeof(mk), hipMemcpyDeviceToHost, st), J3DG_ERR_LAUNCH);
    HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
    const uint32_t m = mk[0], k = mk[1];
    if (m < 2 || k + 1 != m || m > n) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: malformed frontier");
    Tmp fbox, put_idx, put_val;
    if (!alloc(fbox, (size_t)m * 8 * sizeof(float)) || !alloc(put_idx, (size_t)k * sizeof(uint32_t)) ||
        !alloc(put_val, (size_t)k * sizeof(BNode)))
      return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device BVH build: frontier buffers");
    hipLaunchKernelGGL(k_frontier_boxes, dim3((m + 255) / 256), dim3(256), 0, st, (const BNode*)bn,
                       (const float*)box.p, sorted, (const uint32_t*)frontier, m, (float*)fbox.p);
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    std::vector<float> fb(8 * (size_t)m);
    std::vector<uint32_t> top_ids(k);
    HIP_TRY(hipMemcpyAsync(fb.data(), fbox.p, fb.size() * sizeof(float), hipMemcpyDeviceToHost, st), J3DG_ERR_LAUNCH);
    HIP_TRY(hipMemcpyAsync(top_ids.data(), top, k * sizeof(uint32_t), hipMemcpyDeviceToHost, st), J3DG_ERR_LAUNCH);
    HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
    std::vector<Cluster> cl(m);
    for (uint32_t i = 0; i < m; ++i) {
      const float* f = &fb[8 * (size_t)i];
      Cluster& c = cl[i];
      for (int a = 0; a < 3; ++a) c.lo[a] = f[a], c.hi[a] = f[3 + a], c.c[a] = 0.5f * f[a] + 0.5f * f[3 + a];
      std::memcpy(&c.count, f + 6, 4);
      std::memcpy(&c.ref, f + 7, 4);
    }
    std::vector<SahNode> sn;
    sn.reserve(k);
    std::vector<float> right_area(m + 1);
    sah_build(cl, 0, m, sn, right_area);
    if (sn.size() != k) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: top rebuild size");
    // the rebuilt root stays binary node 0; the other rebuilt nodes take the replaced nodes' indices
    std::vector<uint32_t> index_of(k);
    uint32_t next = 0;
    for (uint32_t j = 0; j < k; ++j) {
      if (j == 0) {
        index_of[j] = 0;
        continue;
      }
      while (top_ids[next] == 0) ++next;
      index_of[j] = top_ids[next++];
    }
    std::vector<BNode> vals(k);
    const uint32_t leaf_min = (uint32_t)leaf_size + 1u;
    for (uint32_t j = 0; j < k; ++j) {
      const SahNode& s2 = sn[j];
      BNode& b = vals[j];
      auto resolve = [&](uint32_t r) { return (r & kSahTag) && !(r & kPrimBit) ? index_of[r & ~kSahTag] : r; };
      b.left = resolve(s2.left);
      b.right = resolve(s2.right);
      b.first = 0;
      b.last = std::max(s2.count, leaf_min) - 1u;  // never a leaf: its triangles are not contiguous
      for (int a = 0; a < 3; ++a) b.lo[a] = s2.lo[a], b.hi[a] = s2.hi[a];
    }
    HIP_TRY(hipMemcpyAsync(put_idx.p, index_of.data(), k * sizeof(uint32_t), hipMemcpyHostToDevice, st),
            J3DG_ERR_LAUNCH);
    HIP_TRY(hipMemcpyAsync(put_val.p, vals.data(), k * sizeof(BNode), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
    hipLaunchKernelGGL(k_put_nodes, dim3((k + 255) / 256), dim3(256), 0, st, bn, (const uint32_t*)put_idx.p,
                       (const BNode*)put_val.p, k);
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
  }
  // collapse, level by level from the root (wide node 0 <- binary node 0)
  uint32_t host_c[4] = {1u, 1u, 0u, 0u};
  Item root{0u, 0u, 0u};
  HIP_TRY(hipMemcpyAsync(q0.p, &root, sizeof(Item), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
  uint32_t n_in = 1, levels = 0;
  Item* qin = (Item*)q0.p;
  Item* qout = (Item*)q1.p;
  while (n_in) {
    const uint32_t zero_hdr[4] = {0u, host_c[1], 0u, host_c[3]};
    HIP_TRY(hipMemcpyAsync(cnt, zero_hdr, sizeof(zero_hdr), hipMemcpyHostToDevice, st), J3DG_ERR_LAUNCH);
    CollapseArgs ca{bn, (const float*)box.p, sorted, (uint32_t)leaf_size, qin, n_in, qout, cnt, n,
                    (uint32_t*)wide.p, (uint8_t*)last.p};
    if (width == 8)
      hipLaunchKernelGGL(k_collapse<8>, dim3((n_in + 127) / 128), dim3(128), 0, st, ca);
    else
      hipLaunchKernelGGL(k_collapse<4>, dim3((n_in + 127) / 128), dim3(128), 0, st, ca);
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    HIP_TRY(hipMemcpyAsync(host_c, cnt, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), J3DG_ERR_LAUNCH);
    HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
    if (host_c[2]) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: wide node capacity exceeded");
    ++levels;
    n_in = host_c[0];
    // each level descends at least one binary level, and the radix tree over
    // 63-bit codes with index tie-breaks is at most 92 levels deep
    if (n_in && levels >= 96) return ctx->fail(J3DG_ERR_INTERNAL, "device BVH build: tree deeper than expected");
    std::swap(qin, qout);
  }
  const uint32_t n_wide = host_c[1];
  hipLaunchKernelGGL(k_tri_records, dim3(gb), dim3(256), 0, st, (const float*)m.d_verts, (const uint32_t*)m.d_idx,
                     sorted, (const uint8_t*)last.p, n, m.d_tris);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  const size_t nb = (size_t)n_wide * words * 4;
  if (hipMalloc(&m.d_nodes, nb) != hipSuccess) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device BVH node array");
  HIP_TRY(hipMemcpyAsync(m.d_nodes, wide.p, nb, hipMemcpyDeviceToDevice, st), J3DG_ERR_LAUNCH);
  HIP_TRY(hipStreamSynchronize(st), J3DG_ERR_LAUNCH);
  m.info.n_nodes = n_wide;
  m.info.width = (uint32_t)width;
  m.info.max_depth = levels;
  m.max_stack = host_c[3] + 1;
  m.info.device_bytes = nb + tb + fb + 3 * n_vertices * sizeof(float) + 2 * 3 * (size_t)n * 4 + 6 * sizeof(float4);
  return J3DG_OK;
}

}  // namespace j3d
