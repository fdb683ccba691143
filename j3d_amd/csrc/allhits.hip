// allhits.hip -- the second query of the reference's BVH: every hit of a ray
// in one mesh (jtk::qbvh::find_all_triangles, called at j3d/vox.cpp:333), and
// the voxel grid write_vox builds from it (j3d/vox.cpp:278-376), on the GPU.
//
// One ray per lane, one launch: the lane walks the mesh's BVH with the frame
// kernel's node decode and stack (top entries in LDS, older ones spilled), and
// keeps its hits in a list in global memory sorted by (t, triangle id), the
// first `cap` of them (an insertion per accepted hit).  The result does not
// depend on the visiting order.  For the voxel grid each lane then maps its
// hits to voxels; where several writes land in one voxel, the reference's
// sequential order decides (axis, row, column, hit): each write is a 64-bit
// key (that position, then the palette index) merged with atomicMax, so the
// largest key -- the sequentially last write -- survives.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh_build.h"
#include "j3dg.h"
#include "j3dg_internal.h"
#include "rt_math.h"
#include "traverse.h"

namespace j3d {
namespace {

constexpr int kAllD = 16;             // LDS stack entries per lane
constexpr uint32_t kAllBlock = 256;
constexpr uint32_t kMaxVoxDim = 1024;  // voxel grid side (the key holds 12 bits per row / column index)
constexpr uint32_t kMaxVoxHits = 1u << 24;  // hits per voxel ray (24 bits of the key)

struct HitRec {
  float t;
  uint32_t tri;
  float u, v;
};

struct AllHitsArgs {
  const float4* nodes;
  const float4* tris;
  float inv_cs[16];         // world -> object for query rays
  const float* rays;        // n x 8 world rays; nullptr: the voxel rays of all three axes
  uint32_t n;
  uint32_t axis_first[3];   // voxel rays: first ray of the x, y and z axis (x rays start at 0)
  uint32_t dim[3];
  float bmin[3], bmax[3];
  uint32_t cap;             // stored hits per ray
  HitRec* hits;             // n x cap
  uint32_t* counts;         // nullable: hits per ray (all of them, stored or not)
  uint32_t* max_count;      // voxel rays: largest count over the rays that did not fit
  uint2* spill;
  uint32_t spill_cap;
  unsigned long long* keys;  // voxel grid, 0 = empty
  const float* verts;
  const uint32_t* idx;
  const float* vcol;
  const float* uv;
  const uint32_t* tex;
  uint32_t tex_w, tex_h;
  uint32_t color_mode;      // 0 white, 1 vertex colours, 2 texture
};

__device__ inline bool hit_less(float t, uint32_t id, float t2, uint32_t id2) {
  return t < t2 || (t == t2 && id < id2);
}

// Keeps list[0 .. min(count, cap)) the smallest hits by (t, id), sorted.
__device__ inline void insert_hit(HitRec* list, uint32_t& count, uint32_t cap, float t, uint32_t id, float u,
                                  float v) {
  uint32_t stored = count < cap ? count : cap;
  ++count;
  if (stored == cap) {
    if (cap == 0) return;
    const HitRec last = list[cap - 1];
    if (!hit_less(t, id, last.t, last.tri)) return;
    --stored;
  }
  uint32_t j = stored;
  while (j > 0) {
    const HitRec p = list[j - 1];
    if (!hit_less(t, id, p.t, p.tri)) break;
    list[j] = p;
    --j;
  }
  list[j] = HitRec{t, id, u, v};
}

// (uint32_t)x for a coordinate that should lie in [0, n): truncation toward
// zero, saturated into [0, n - 1], NaN -> 0 (the reference's casts are
// undefined outside the range).
__device__ inline uint32_t vox_index(float x, uint32_t n) {
  if (!(x > 0.f)) return 0u;
  if (x >= (float)n) return n - 1u;
  const uint32_t i = (uint32_t)x;
  return i < n ? i : n - 1u;
}

// color_to_index (j3d/vox.cpp:154-171)
__device__ inline uint32_t vox_color_index(uint32_t r, uint32_t g, uint32_t b) {
  r = r < 16u ? 0u : r - 16u;
  g = g < 16u ? 0u : g - 16u;
  b = b < 32u ? 0u : b - 32u;
  const uint32_t ret = ((r >> 5) << 5) | ((g >> 5) << 2) | (b >> 6);
  return ret ? ret : 1u;
}

template <int W>
__global__ void __launch_bounds__(kAllBlock) k_all_hits(AllHitsArgs a) {
  extern __shared__ uint2 lds_stack[];
  Stack<kAllD> st;
  st.init(lds_stack, a.spill, a.spill_cap);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const bool vox = a.rays == nullptr;
  uint32_t d1 = 0, d2 = 0, ax = 0;
  Trav s;
  bool finite = true;
  if (!vox) {
    const float* q = a.rays + 8 * (size_t)i;
    const f4 o = {q[0], q[1], q[2], 1.f}, d = {q[3], q[4], q[5], 0.f};
    finite = ray_finite(o, d);
    s.r = make_object_ray(a.inv_cs, o, d, q[6], q[7]);
  } else {
    // write_vox's ray of row d1, column d2 along axis ax (j3d/vox.cpp:317-330)
    ax = i >= a.axis_first[2] ? 2u : (i >= a.axis_first[1] ? 1u : 0u);
    const uint32_t ax1 = (ax + 1) % 3, ax2 = (ax + 2) % 3, r = i - a.axis_first[ax];
    d1 = r / a.dim[ax2];
    d2 = r % a.dim[ax2];
    float o[3], d[3] = {0.f, 0.f, 0.f};
    o[ax] = a.bmin[ax];
    o[ax1] = (((float)d1 + 0.5f) / (float)a.dim[ax1]) * (a.bmax[ax1] - a.bmin[ax1]) + a.bmin[ax1];
    o[ax2] = (((float)d2 + 0.5f) / (float)a.dim[ax2]) * (a.bmax[ax2] - a.bmin[ax2]) + a.bmin[ax2];
    d[ax] = ax == 2 ? 2.f : 1.f;
    const f4 oo = {o[0], o[1], o[2], 1.f}, dd = {d[0], d[1], d[2], 0.f};
    finite = ray_finite(oo, dd);  // a NaN or infinite mesh box gives rays that hit nothing
    s.r = world_ray(oo, dd, 0.f, FLT_MAX);
  }
  s.nodes = a.nodes;
  s.tris = a.tris;
  s.thi = s.r.t_far * kSlackUp;
  s.tmax = s.r.t_far;
  s.cur = finite ? 0u : kNone;  // the root, an internal node; a ray that is not finite hits nothing
  s.obj = 0;
  s.base = 0;
  s.any = false;
  Counters c = {0, 0, 0};
  HitRec* list = a.hits + (size_t)i * a.cap;
  uint32_t count = 0;
  for (;;) {
    if (s.cur != kNone) {
      if (s.cur & kLeafBit) {
        const uint32_t first = s.cur & kLeafIndexMask;
        const uint32_t cnt = ((s.cur >> kLeafCountShift) & 3u) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
          const float4* tp = s.tris + (size_t)(first + k) * 3;
          const float4 v0 = gload4(tp), e1 = gload4(tp + 1), e2 = gload4(tp + 2);
          float t, u, v;
          if (intersect_triangle(s.r.ox, s.r.oy, s.r.oz, s.r.dx, s.r.dy, s.r.dz, v0.x, v0.y, v0.z, e1.x, e1.y, e1.z,
                                 e2.x, e2.y, e2.z, s.r.t_near, s.r.t_far, t, u, v))
            insert_hit(list, count, a.cap, t, __float_as_uint(v0.w), u, v);
        }
        s.cur = kNone;
      } else {
        s.cur = visit_node<W, kAllD, false>(s.nodes, s.cur, s.r, s.thi, st, c);
      }
    }
    if (s.cur == kNone && pop_next<kAllD>(s, st)) break;
  }
  if (a.counts) a.counts[i] = count;
  if (!vox) return;
  if (count > a.cap) {  // the host grows the lists and runs the launch again
    atomicMax(a.max_count, count);
    return;
  }
  // every hit to its voxel (j3d/vox.cpp:335-375)
  for (uint32_t k = 0; k < count; ++k) {
    const HitRec h = list[k];
    const uint32_t* tr = a.idx + 3 * (size_t)h.tri;
    const float* p0 = a.verts + 3 * (size_t)tr[0];
    const float* p1 = a.verts + 3 * (size_t)tr[1];
    const float* p2 = a.verts + 3 * (size_t)tr[2];
    const float w = 1.f - h.u - h.v;
    float pos[3], clr[3] = {1.f, 1.f, 1.f};
    for (int j = 0; j < 3; ++j) pos[j] = p0[j] * w + h.u * p1[j] + h.v * p2[j];
    if (a.color_mode == 2) {
      const float* uvc = a.uv + 6 * (size_t)h.tri;
      float cc[2];
      for (int j = 0; j < 2; ++j) {
        cc[j] = w * uvc[j] + h.u * uvc[2 + j] + h.v * uvc[4 + j];
        // std::max(std::min(c, 1.f), 0.f), NaN kept (vox.cpp:345-346)
        const float lo = 1.f < cc[j] ? 1.f : cc[j];
        cc[j] = lo < 0.f ? 0.f : lo;
      }
      const uint32_t x = vox_index(cc[0] * (float)(a.tex_w - 1), a.tex_w);
      const uint32_t y = vox_index(cc[1] * (float)(a.tex_h - 1), a.tex_h);
      const uint32_t col = a.tex[(size_t)y * a.tex_w + x];
      clr[0] = (float)(col & 255u) / 255.f;
      clr[1] = (float)((col >> 8) & 255u) / 255.f;
      clr[2] = (float)((col >> 16) & 255u) / 255.f;
    } else if (a.color_mode == 1) {
      const float* c0 = a.vcol + 3 * (size_t)tr[0];
      const float* c1 = a.vcol + 3 * (size_t)tr[1];
      const float* c2 = a.vcol + 3 * (size_t)tr[2];
      for (int j = 0; j < 3; ++j) clr[j] = c0[j] * w + h.u * c1[j] + h.v * c2[j];
    }
    uint32_t X[3];
    for (int j = 0; j < 3; ++j)
      X[j] = vox_index((pos[j] - a.bmin[j]) / (a.bmax[j] - a.bmin[j]) * (float)a.dim[j], a.dim[j]);
    const uint32_t ci = vox_color_index(color_byte(clr[0]), color_byte(clr[1]), color_byte(clr[2]));
    unsigned long long key = ax;
    key = (key << 12) | d1;
    key = (key << 12) | d2;
    key = (key << 24) | k;
    key = (key << 8) | ci;
    atomicMax(a.keys + X[0] + ((size_t)X[1] + (size_t)X[2] * a.dim[1]) * a.dim[0], key);
  }
}

__global__ void k_vox_final(const unsigned long long* __restrict__ keys, uint64_t n, uint8_t* __restrict__ data) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    data[i] = (uint8_t)(keys[i] & 255ull);
}

// write_vox's grid size (j3d/vox.cpp:284-297), in float: dim_k = max_dim *
// extent_k / largest extent, at least 1; a NaN ratio counts as 0.
void vox_dims(const float* mn, const float* mx, uint32_t max_dim, uint32_t dim[3]) {
  int L = 0;
  if ((mx[1] - mn[1]) > (mx[L] - mn[L])) L = 1;
  if ((mx[2] - mn[2]) > (mx[L] - mn[L])) L = 2;
  for (int j = 0; j < 3; ++j) {
    const float q = (float)max_dim * (mx[j] - mn[j]) / (mx[L] - mn[L]);
    uint32_t d = 0;
    if (q > 0.f) d = q >= (float)max_dim ? max_dim : (uint32_t)q;
    dim[j] = d ? d : 1u;
  }
}

void fill_mesh_args(const Mesh& m, AllHitsArgs& a) {
  std::memset(&a, 0, sizeof(a));
  a.nodes = m.d_nodes;
  a.tris = m.d_tris;
  std::memcpy(a.inv_cs, m.inv_cs, sizeof(a.inv_cs));
  a.verts = m.d_verts;
  a.idx = m.d_idx;
}

int launch_all_hits(Ctx* ctx, const Mesh& m, AllHitsArgs& a) {
  const uint32_t grid = (a.n + kAllBlock - 1) / kAllBlock;
  const uint32_t cap = m.max_stack > (uint32_t)kAllD ? m.max_stack - kAllD + 1 : 1;
  int rc = ctx->ensure_spill((uint64_t)grid * kAllBlock, cap);
  if (rc) return rc;
  a.spill = ctx->spill;
  a.spill_cap = cap;
  const size_t lds = (size_t)kAllD * kAllBlock * sizeof(uint2);
  if (m.info.width == 8)
    hipLaunchKernelGGL(k_all_hits<8>, dim3(grid), dim3(kAllBlock), lds, ctx->stream, a);
  else
    hipLaunchKernelGGL(k_all_hits<4>, dim3(grid), dim3(kAllBlock), lds, ctx->stream, a);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // namespace
}  // namespace j3d

using j3d::Ctx;
using j3d::Mesh;

extern "C" {

int j3dg_all_hits(j3dg_ctx* h, uint32_t slot, const float* rays, uint64_t n, uint32_t max_hits, uint32_t* counts,
                  uint32_t* tri, float* t, float* u, float* v) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if ((!rays || !counts) && n) return ctx->fail(J3DG_ERR_INVALID_ARG, "null rays or counts");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  if (n > (1ull << 31)) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many rays in one call");
  if (max_hits > 65536 || n * (uint64_t)max_hits > (1ull << 28))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "n * max_hits exceeds 2^28 stored hits");
  if (n == 0) return J3DG_OK;
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  const Mesh& m = ctx->meshes[slot];
  void* d_rays = ctx->scratch_buffer(0, n * 8 * sizeof(float));
  void* d_hits = ctx->scratch_buffer(1, n * max_hits * sizeof(j3d::HitRec));
  void* d_counts = ctx->scratch_buffer(2, n * sizeof(uint32_t));
  if (!d_rays || !d_hits || !d_counts) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "all-hits buffers");
  HIP_CALL(hipMemcpyAsync(d_rays, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice, ctx->stream),
           J3DG_ERR_LAUNCH);
  j3d::AllHitsArgs a;
  j3d::fill_mesh_args(m, a);
  a.rays = static_cast<const float*>(d_rays);
  a.n = (uint32_t)n;
  a.cap = max_hits;
  a.hits = static_cast<j3d::HitRec*>(d_hits);
  a.counts = static_cast<uint32_t*>(d_counts);
  int rc = j3d::launch_all_hits(ctx, m, a);
  if (rc) return rc;
  HIP_CALL(hipMemcpyAsync(counts, d_counts, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
           J3DG_ERR_LAUNCH);
  std::vector<j3d::HitRec> hits((size_t)n * max_hits);
  if (max_hits)
    HIP_CALL(hipMemcpyAsync(hits.data(), d_hits, hits.size() * sizeof(j3d::HitRec), hipMemcpyDeviceToHost,
                            ctx->stream),
             J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t k = 0; k < max_hits; ++k) {
      const size_t o = (size_t)i * max_hits + k;
      const bool has = k < counts[i];
      if (tri) tri[o] = has ? hits[o].tri : 0xFFFFFFFFu;
      if (t) t[o] = has ? hits[o].t : 0.f;
      if (u) u[o] = has ? hits[o].u : 0.f;
      if (v) v[o] = has ? hits[o].v : 0.f;
    }
  return J3DG_OK;
}

int j3dg_voxelize(j3dg_ctx* h, uint32_t slot, uint32_t max_dim, uint32_t dim[3], uint8_t* data, uint64_t capacity) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!dim) return ctx->fail(J3DG_ERR_INVALID_ARG, "null dim");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  if (max_dim == 0 || max_dim > j3d::kMaxVoxDim) return ctx->fail(J3DG_ERR_INVALID_ARG, "max_dim must be 1 .. 1024");
  const Mesh& m = ctx->meshes[slot];
  j3d::vox_dims(m.info.bb_min, m.info.bb_max, max_dim, dim);
  const uint64_t nvox = (uint64_t)dim[0] * dim[1] * dim[2];
  if (!data) return J3DG_OK;
  if (capacity < nvox) return ctx->fail(J3DG_ERR_INVALID_ARG, "voxel buffer smaller than dim[0] * dim[1] * dim[2]");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  // scratch 0: keys, then the uint8 grid after them; 1: hit lists; 2: the overflow count
  void* d_keys = ctx->scratch_buffer(0, nvox * sizeof(unsigned long long) + nvox);
  void* d_max = ctx->scratch_buffer(2, sizeof(uint32_t));
  if (!d_keys || !d_max) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "voxel buffers");
  uint8_t* d_data = static_cast<uint8_t*>(d_keys) + nvox * sizeof(unsigned long long);
  HIP_CALL(hipMemsetAsync(d_keys, 0, nvox * sizeof(unsigned long long), ctx->stream), J3DG_ERR_LAUNCH);
  j3d::AllHitsArgs a;
  j3d::fill_mesh_args(m, a);
  for (int k = 0; k < 3; ++k) a.dim[k] = dim[k], a.bmin[k] = m.info.bb_min[k], a.bmax[k] = m.info.bb_max[k];
  a.keys = static_cast<unsigned long long*>(d_keys);
  a.max_count = static_cast<uint32_t*>(d_max);
  // texture over vertex colours over white (j3d/vox.cpp:301-304)
  if (m.d_uv && m.d_tex && m.tex_w && m.tex_h) {
    a.color_mode = 2, a.uv = m.d_uv, a.tex = m.d_tex, a.tex_w = m.tex_w, a.tex_h = m.tex_h;
  } else if (m.d_vcol) {
    a.color_mode = 1, a.vcol = m.d_vcol;
  }
  // the three axes' rays in one launch: x rays, then y, then z
  a.axis_first[0] = 0;
  a.axis_first[1] = dim[1] * dim[2];
  a.axis_first[2] = a.axis_first[1] + dim[2] * dim[0];
  const uint32_t n = a.axis_first[2] + dim[0] * dim[1];
  uint32_t cap = 32;
  for (;;) {
    void* d_hits = ctx->scratch_buffer(1, (size_t)n * cap * sizeof(j3d::HitRec));
    if (!d_hits) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "voxel hit lists");
    HIP_CALL(hipMemsetAsync(d_max, 0, sizeof(uint32_t), ctx->stream), J3DG_ERR_LAUNCH);
    a.n = n;
    a.cap = cap;
    a.hits = static_cast<j3d::HitRec*>(d_hits);
    int rc = j3d::launch_all_hits(ctx, m, a);
    if (rc) return rc;
    uint32_t mc = 0;
    HIP_CALL(hipMemcpyAsync(&mc, d_max, sizeof(mc), hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
    HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
    if (mc <= cap) break;
    if (mc > j3d::kMaxVoxHits) return ctx->fail(J3DG_ERR_UNSUPPORTED, "more than 2^24 hits on one voxel ray");
    while (cap < mc) cap *= 2;  // the rays that fitted wrote their voxels already; writing them again is harmless
  }
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((nvox + 255) / 256, 65536);
  hipLaunchKernelGGL(j3d::k_vox_final, dim3(blocks), dim3(256), 0, ctx->stream,
                     (const unsigned long long*)d_keys, nvox, d_data);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpyAsync(data, d_data, nvox, hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // extern "C"
