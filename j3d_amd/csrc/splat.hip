// splat.hip -- point clouds drawn over a rendered target
// (canvas::render_pointclouds_on_image, j3d/canvas.cpp:952-1030).
//
// The reference draws the clouds one after another, each point in order, into
// one pixel under a z-buffer of 1/depth filled from the G-buffer
// (canvas.cpp:963-970); a point is drawn when its 1/w is greater than the
// z-buffer value, and then writes object_id = its index, depth = 1/(1/w) and
// the cloud's db_id into the G-buffer (canvas.cpp:1000-1026) and its colour
// into the image.  What survives in a pixel is the point with the greatest
// 1/w, the first of them on a tie, or nothing when the z-buffer value is not
// beaten.
//
// Here each pixel holds a 64-bit key: the order-preserving bits of its
// z-buffer value, then 0xFFFFFFFF for what the pixel already shows or
// 0xFFFFFFFE - i for point i.  Per cloud one launch over the points merges
// keys with atomicMax (the greatest 1/w wins; on a tie the lower index, and
// what the pixel already shows beats every point), then one launch over the
// pixels writes the winner's fields and marks the key as shown, so that the
// next cloud's points lose ties to it.  The result does not depend on the
// order in which the points are processed; a point first reads its pixel's
// key and skips the atomic when the key is already at least its own.
#include <hip/hip_runtime.h>

#include <cstring>

#include "j3dg.h"
#include "j3dg_internal.h"
#include "rt_math.h"

namespace j3d {
namespace {

constexpr uint32_t kSplatBlock = 256;
constexpr int kSplatBatch = 4;  // points per thread and loop trip in k_splat_depth
constexpr uint64_t kMaxCloudPoints = 0xFFFFFFFEull;  // point i's key holds 0xFFFFFFFE - i

struct SplatMats {
  float cs[16], cs_inv[16], proj[16];
};

// Order-preserving image of a float's bits (larger float -> larger key); NaN
// maps above +inf, so that a NaN z-buffer value is never beaten, as a
// comparison with NaN is false.
__device__ inline uint32_t order_bits(float f) {
  if (f != f) return 0xFFFFFFFFu;
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ inline float order_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ void k_splat_init(const j3dg_pixel* __restrict__ gbuf, unsigned long long* __restrict__ keys,
                             uint32_t n_pix) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pix) return;
  const j3dg_pixel& p = gbuf[i];
  const float z = p.db_id != 0 ? 1.f / p.depth : 0.f;
  keys[i] = ((unsigned long long)order_bits(z) << 32) | 0xFFFFFFFFull;
}

__global__ void __launch_bounds__(kSplatBlock) k_splat_depth(const float* __restrict__ xyz, uint64_t n, SplatMats m,
                                                             uint32_t w, uint32_t h,
                                                             unsigned long long* __restrict__ keys) {
  // kSplatBatch points per thread and trip (i, i + stride, ...): their point
  // and key reads are issued together, then the atomics, so that the reads of
  // one trip overlap instead of waiting behind the previous point's atomic.
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; base < n; base += stride * kSplatBatch) {
    uint32_t pixel[kSplatBatch];
    unsigned long long key[kSplatBatch], seen[kSplatBatch];
    bool ok[kSplatBatch];
#pragma unroll
    for (int b = 0; b < kSplatBatch; ++b) {
      const uint64_t i = base + (uint64_t)b * stride;
      float inv_w = 0.f;
      pixel[b] = 0;
      ok[b] = i < n &&
              splat_project(m.cs, m.cs_inv, m.proj, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], w, h, pixel[b], inv_w);
      key[b] = ((unsigned long long)order_bits(inv_w) << 32) | (unsigned long long)(0xFFFFFFFEu - (uint32_t)i);
    }
    // Keys only grow during this launch, so any value read here (untorn: a
    // 64-bit atomic load, possibly older than the latest) is a lower bound of
    // the final key: when it is already >= this key the atomic cannot change
    // anything and is skipped.  Most points lose, so most atomics go.
#pragma unroll
    for (int b = 0; b < kSplatBatch; ++b)
      seen[b] = ok[b] ? __hip_atomic_load(keys + pixel[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : ~0ull;
#pragma unroll
    for (int b = 0; b < kSplatBatch; ++b)
      if (ok[b] && key[b] > seen[b]) atomicMax(keys + pixel[b], key[b]);
  }
}

__global__ void k_splat_resolve(unsigned long long* __restrict__ keys, uint32_t n_pix, const float* __restrict__ normals,
                                const uint32_t* __restrict__ colors, SplatMats m, bool shade, uint32_t db_id,
                                j3dg_pixel* __restrict__ gbuf, uint32_t* __restrict__ rgba) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pix) return;
  const unsigned long long key = keys[i];
  const uint32_t low = (uint32_t)key;
  if (low == 0xFFFFFFFFu) return;
  const uint32_t v = 0xFFFFFFFEu - low;
  const float inv_w = order_float((uint32_t)(key >> 32));
  j3dg_pixel& p = gbuf[i];
  p.object_id = v;
  p.depth = 1.f / inv_w;
  p.db_id = db_id;
  const uint32_t c = colors ? colors[v] : 0xffffffffu;
  rgba[i] = normals ? splat_color(c, shade, m.cs, m.cs_inv, normals[3 * (size_t)v], normals[3 * (size_t)v + 1],
                                  normals[3 * (size_t)v + 2])
                    : c;
  keys[i] = key | 0xFFFFFFFFull;
}

uint32_t splat_grid(uint64_t n, int n_cus) {
  const uint64_t blocks = (n + kSplatBlock - 1) / kSplatBlock;
  const uint64_t cap = (uint64_t)n_cus * 32;
  return (uint32_t)std::max<uint64_t>(1, std::min(blocks, cap));
}

}  // namespace
}  // namespace j3d

using j3d::Ctx;
using j3d::PointCloud;

extern "C" {

int j3dg_upload_points(j3dg_ctx* h, uint32_t db_id, const float* xyz, uint64_t n_points, const float* normals,
                       const uint32_t* colors, const float cs[16], uint32_t* cloud) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!cloud || (!xyz && n_points)) return ctx->fail(J3DG_ERR_INVALID_ARG, "null points or cloud");
  if (n_points > j3d::kMaxCloudPoints) return ctx->fail(J3DG_ERR_INVALID_ARG, "more than 2^32 - 2 points");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  PointCloud c;
  c.live = true;
  c.db_id = db_id;
  c.n = n_points;
  for (int i = 0; i < 16; ++i) c.cs[i] = cs ? cs[i] : (i % 5 == 0 ? 1.f : 0.f);
  const size_t b3 = (size_t)n_points * 3 * sizeof(float);
  auto up = [&](void** dst, const void* src, size_t bytes) -> bool {
    if (!src || !bytes) return true;
    if (hipMalloc(dst, bytes) != hipSuccess) {
      *dst = nullptr;
      return false;
    }
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void**)&c.d_xyz, xyz, b3) || !up((void**)&c.d_normals, normals, b3) ||
      !up((void**)&c.d_colors, colors, (size_t)n_points * sizeof(uint32_t))) {
    c.release();
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "point cloud upload");
  }
  uint32_t slot = 0;
  while (slot < ctx->clouds.size() && ctx->clouds[slot].live) ++slot;
  if (slot == ctx->clouds.size()) ctx->clouds.push_back(c);
  else ctx->clouds[slot] = c;
  *cloud = slot;
  return J3DG_OK;
}

int j3dg_remove_points(j3dg_ctx* h, uint32_t cloud) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (cloud >= ctx->clouds.size() || !ctx->clouds[cloud].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown cloud");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  ctx->clouds[cloud].release();
  return J3DG_OK;
}

int j3dg_set_points_cs(j3dg_ctx* h, uint32_t cloud, const float cs[16]) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!cs) return ctx->fail(J3DG_ERR_INVALID_ARG, "null cs");
  if (cloud >= ctx->clouds.size() || !ctx->clouds[cloud].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown cloud");
  std::memcpy(ctx->clouds[cloud].cs, cs, sizeof(float) * 16);
  return J3DG_OK;
}

int j3dg_splat_points(j3dg_ctx* h, uint32_t target, const float proj[16], const float cs_inv[16], uint32_t flags,
                      const uint32_t* clouds, uint32_t n_clouds) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!proj || !cs_inv || (!clouds && n_clouds)) return ctx->fail(J3DG_ERR_INVALID_ARG, "null argument");
  if (target >= ctx->targets.size() || !ctx->targets[target].live)
    return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  for (uint32_t k = 0; k < n_clouds; ++k)
    if (clouds[k] >= ctx->clouds.size() || !ctx->clouds[clouds[k]].live)
      return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown cloud");
  if (n_clouds == 0) return J3DG_OK;  // canvas.cpp:956: nothing to draw
  j3d::Target& tg = ctx->targets[target];
  const uint64_t n_pix = (uint64_t)tg.w * tg.h;
  if (n_pix == 0) return J3DG_OK;
  if (n_pix >= (1ull << 32)) return ctx->fail(J3DG_ERR_INVALID_ARG, "target too large");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  auto* keys = static_cast<unsigned long long*>(ctx->scratch_buffer(3, n_pix * sizeof(unsigned long long)));
  if (!keys) return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "splat z-buffer");
  const uint32_t pix_blocks = (uint32_t)((n_pix + j3d::kSplatBlock - 1) / j3d::kSplatBlock);
  hipLaunchKernelGGL(j3d::k_splat_init, dim3(pix_blocks), dim3(j3d::kSplatBlock), 0, ctx->stream, tg.gbuf, keys,
                     (uint32_t)n_pix);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  j3d::SplatMats m;
  std::memcpy(m.cs_inv, cs_inv, sizeof(m.cs_inv));
  std::memcpy(m.proj, proj, sizeof(m.proj));
  const bool one_bit = (flags & J3DG_ONE_BIT) != 0;
  const bool shade = (flags & J3DG_SHADING) != 0;
  ctx->splat_ms = 0.f;
  for (uint32_t k = 0; k < n_clouds; ++k) {
    const PointCloud& c = ctx->clouds[clouds[k]];
    if (c.n == 0) continue;
    std::memcpy(m.cs, c.cs, sizeof(m.cs));
    HIP_CALL(hipEventRecord(ctx->splat_ev[0], ctx->stream), J3DG_ERR_LAUNCH);
    hipLaunchKernelGGL(j3d::k_splat_depth, dim3(j3d::splat_grid(c.n, ctx->n_cus)), dim3(j3d::kSplatBlock), 0,
                       ctx->stream, c.d_xyz, c.n, m, tg.w, tg.h, keys);
    HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
    HIP_CALL(hipEventRecord(ctx->splat_ev[1], ctx->stream), J3DG_ERR_LAUNCH);
    // canvas.cpp:994-995: normals only with shading, colours not in one-bit mode
    hipLaunchKernelGGL(j3d::k_splat_resolve, dim3(pix_blocks), dim3(j3d::kSplatBlock), 0, ctx->stream, keys,
                       (uint32_t)n_pix, shade ? c.d_normals : nullptr, one_bit ? nullptr : c.d_colors, m, shade,
                       c.db_id, tg.gbuf, tg.rgba);
    HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
    float ms = 0.f;
    HIP_CALL(hipEventSynchronize(ctx->splat_ev[1]), J3DG_ERR_LAUNCH);
    if (hipEventElapsedTime(&ms, ctx->splat_ev[0], ctx->splat_ev[1]) == hipSuccess) ctx->splat_ms += ms;
  }
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // extern "C"
