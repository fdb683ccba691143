// j3dg.hip -- the gfx950 ray-casting path of j3d behind the C ABI in include/j3dg.h.
//
// Per frame (canvas::render_scene + canvas::canvas_to_image, j3d/canvas.cpp:582-898):
//   k_trace  : primary ray generation, closest-hit traversal of each scene
//              object's wide BVH, optional shadow ray, G-buffer write
//   k_shade  : background gradient + matcap/edge shading of the G-buffer
// Compiled with -ffp-contract=off: ray generation, the triangle test, normals
// and shading round exactly as the host-side restatement does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bvh_build.h"
#include "j3dg.h"
#include "j3dg_internal.h"
#include "rt_math.h"

using namespace j3d;

// ============================================================ device side ===

namespace {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kSlackUp = 1.0f + 1.0f / 65536.0f;    // t comparisons: 2^-16 relative slack
constexpr float kSlackDown = 1.0f - 1.0f / 65536.0f;

struct Hit {
  float t, u, v;
  uint32_t tri, obj;
  bool found;
};

struct Ray {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz;  // safe reciprocal direction
  float t_near, t_far;
};

__device__ inline float safe_rcp(float d) {
  float a = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
  return 1.f / a;
}

__device__ inline Ray make_object_ray(const float* inv_cs, f4 o, f4 d, float tn, float tf) {
  f4 oo = mat_vec(inv_cs, o);
  f4 dd = mat_vec(inv_cs, d);
  Ray r;
  r.ox = oo.x, r.oy = oo.y, r.oz = oo.z;
  r.dx = dd.x, r.dy = dd.y, r.dz = dd.z;
  r.ix = safe_rcp(dd.x), r.iy = safe_rcp(dd.y), r.iz = safe_rcp(dd.z);
  r.t_near = tn;
  r.t_far = tf;
  return r;
}

// Per-lane traversal stack: the top `D` entries live in LDS (interleaved by
// lane, conflict-free), older entries spill to a global scratch area laid out
// [entry][lane] so that spills of one wave coalesce.
template <int D>
struct Stack {
  uint2* lds;         // &lds[0][tid], stride `stride`
  uint32_t stride;    // threads per block
  uint2* spill;       // &spill[0][global lane]
  uint64_t sstride;   // total lanes
  uint32_t cap;       // spill entries available
  uint32_t sp;
  bool overflow;

  __device__ void push(uint32_t ref, float dist) {
    uint32_t slot = sp & (D - 1);
    if (sp >= (uint32_t)D) {
      uint32_t e = sp - D;
      if (e >= cap) {  // bounded by the builder's max_stack; never expected
        overflow = true;
        return;
      }
      spill[(uint64_t)e * sstride] = lds[slot * stride];
    }
    lds[slot * stride] = make_uint2(ref, __float_as_uint(dist));
    ++sp;
  }
  __device__ bool pop(uint32_t& ref, float& dist) {
    if (sp == 0) return false;
    --sp;
    uint32_t slot = sp & (D - 1);
    uint2 v = lds[slot * stride];
    if (sp >= (uint32_t)D) lds[slot * stride] = spill[(uint64_t)(sp - D) * sstride];
    ref = v.x;
    dist = __uint_as_float(v.y);
    return true;
  }
};

struct Counters {
  uint64_t nodes, tris;
};

// Visits one wide node: pushes the hit children far-to-near, returns the nearest
// (or kNone).  Culling is conservative (see the slack constants).
template <int W, int D, bool COUNT>
__device__ inline uint32_t visit_node(const float4* __restrict__ nodes, uint32_t node, const Ray& r, float thi,
                                      Stack<D>& st, Counters& c) {
  const float4* np = nodes + (size_t)node * (2 * W);
  float b[7 * W];
#pragma unroll
  for (int i = 0; i < 7 * W / 4; ++i) {
    float4 q = np[i];
    b[4 * i] = q.x, b[4 * i + 1] = q.y, b[4 * i + 2] = q.z, b[4 * i + 3] = q.w;
  }
  if (COUNT) c.nodes++;
  const float tlo = r.t_near * kSlackDown;
  float dist[W];
  uint32_t ref[W];
  const bool sx = r.ix < 0.f, sy = r.iy < 0.f, sz = r.iz < 0.f;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    float nx = sx ? b[1 * W + i] : b[0 * W + i], fx = sx ? b[0 * W + i] : b[1 * W + i];
    float ny = sy ? b[3 * W + i] : b[2 * W + i], fy = sy ? b[2 * W + i] : b[3 * W + i];
    float nz = sz ? b[5 * W + i] : b[4 * W + i], fz = sz ? b[4 * W + i] : b[5 * W + i];
    float tnx = (nx - r.ox) * r.ix, tfx = (fx - r.ox) * r.ix;
    float tny = (ny - r.oy) * r.iy, tfy = (fy - r.oy) * r.iy;
    float tnz = (nz - r.oz) * r.iz, tfz = (fz - r.oz) * r.iz;
    float en = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tlo));
    float ex = fminf(fminf(fminf(tfx, tfy), tfz) * kSlackUp, thi);
    dist[i] = en <= ex ? en : INFINITY;
    ref[i] = __float_as_uint(b[6 * W + i]);
  }
  // sorting network, ascending distance
#define J3D_CSWAP(a, bb)                                      \
  if (dist[bb] < dist[a]) {                                    \
    float td = dist[a];                                        \
    dist[a] = dist[bb];                                        \
    dist[bb] = td;                                             \
    uint32_t tr = ref[a];                                      \
    ref[a] = ref[bb];                                          \
    ref[bb] = tr;                                              \
  }
  if (W == 4) {
    J3D_CSWAP(0, 1) J3D_CSWAP(2, 3) J3D_CSWAP(0, 2) J3D_CSWAP(1, 3) J3D_CSWAP(1, 2)
  } else {
    J3D_CSWAP(0, 2) J3D_CSWAP(1, 3) J3D_CSWAP(4, 6) J3D_CSWAP(5, 7)
    J3D_CSWAP(0, 4) J3D_CSWAP(1, 5) J3D_CSWAP(2, 6) J3D_CSWAP(3, 7)
    J3D_CSWAP(0, 1) J3D_CSWAP(2, 3) J3D_CSWAP(4, 5) J3D_CSWAP(6, 7)
    J3D_CSWAP(2, 4) J3D_CSWAP(3, 5)
    J3D_CSWAP(1, 4) J3D_CSWAP(3, 6)
    J3D_CSWAP(1, 2) J3D_CSWAP(3, 4) J3D_CSWAP(5, 6)
  }
#undef J3D_CSWAP
#pragma unroll
  for (int i = W - 1; i >= 1; --i)
    if (dist[i] != INFINITY) st.push(ref[i], dist[i]);
  return dist[0] != INFINITY ? ref[0] : kNone;
}

// Closest hit (or any hit) inside one object.  `best` carries the running
// lexicographic minimum of (t, object index, triangle id) across objects.
template <int W, int D, bool ANY, bool COUNT>
__device__ bool trace_object(const float4* __restrict__ nodes, const float4* __restrict__ tris, uint32_t oi,
                             const Ray& r, Hit& best, Stack<D>& st, Counters& c) {
  float tmax = best.found ? fminf(best.t, r.t_far) : r.t_far;
  float thi = tmax * kSlackUp;
  uint32_t cur = 0;  // root is an internal node
  st.sp = 0;
  for (;;) {
    if (cur == kNone) {
      uint32_t ref;
      float dist;
      bool got = false;
      while (st.pop(ref, dist)) {
        if (dist <= thi) {
          got = true;
          break;
        }
      }
      if (!got) break;
      cur = ref;
    }
    if (cur & kLeafBit) {
      uint32_t k = cur & ~kLeafBit;
      for (;;) {
        const float4* tp = tris + (size_t)k * 3;
        float4 a = tp[0], e1 = tp[1], e2 = tp[2];
        if (COUNT) c.tris++;
        float t, u, v;
        if (intersect_triangle(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, a.x, a.y, a.z, e1.x, e1.y, e1.z, e2.x, e2.y,
                               e2.z, r.t_near, tmax, t, u, v)) {
          if (ANY) return true;
          uint32_t id = __float_as_uint(a.w);
          bool better = !best.found || t < best.t || (t == best.t && (oi < best.obj || (oi == best.obj && id < best.tri)));
          if (better) {
            best.found = true;
            best.t = t, best.u = u, best.v = v, best.tri = id, best.obj = oi;
            tmax = t;
            thi = t * kSlackUp;
          }
        }
        if (__float_as_uint(e1.w) & 1u) break;
        ++k;
      }
      cur = kNone;
    } else {
      cur = visit_node<W, D, COUNT>(nodes, cur, r, thi, st, c);
    }
  }
  return false;
}

}  // namespace

namespace j3d {

// Primary-ray tile of 8x8 pixels per wave.
__device__ inline bool tile_pixel(const FrameParams& f, uint32_t tile, uint32_t lane, uint32_t& px, uint32_t& py) {
  uint32_t tx = tile % f.tiles_x, ty = tile / f.tiles_x;
  px = (uint32_t)f.x0 + tx * 8 + (lane & 7);
  py = (uint32_t)f.y0 + ty * 8 + (lane >> 3);
  return px <= (uint32_t)f.x1 && py <= (uint32_t)f.y1;
}

template <int W, int D, bool COUNT>
__device__ void trace_pixel(const FrameParams& f, const DevObject* __restrict__ objs, uint32_t px, uint32_t py,
                            Stack<D>& st, Counters& c, uint32_t* lstats) {
  const float Wf = (float)f.w, Hf = (float)f.h;
  // j3d/canvas.cpp:773-782
  f4 sp = {2.f * (((float)px + 0.5f) / Wf) - 1.f, 2.f * (((float)py + 0.5f) / Hf) - 1.f, f.near_plane, 1.f};
  f4 dir = mat_vec(f.proj_inv, sp);
  dir.w = 0.f;
  dir = mat_vec(f.cs, dir);
  const f4 origin = mat_vec(f.cs, f4{0.f, 0.f, 0.f, 1.f});
  const float t_near = f.diagonal / 100.f;
  Hit best;
  best.found = false;
  best.t = 0.f, best.u = 0.f, best.v = 0.f, best.tri = kMiss, best.obj = kMiss;
  for (uint32_t i = 0; i < f.n_obj; ++i) {
    const DevObject& ob = objs[i];
    Ray r = make_object_ray(ob.inv_cs, origin, dir, t_near, FLT_MAX);
    trace_object<W, D, false, COUNT>(ob.nodes, ob.tris, i, r, best, st, c);
  }
  lstats[0]++;
  char* pix = reinterpret_cast<char*>(f.gbuf) + ((size_t)py * f.w + px) * 32;
  if (best.found) {
    lstats[1]++;
    const DevObject& ob = objs[best.obj];
    // j3d/canvas.cpp:790-801
    const float* tn = ob.normals + 3 * (size_t)best.tri;
    f4 n = {tn[0], tn[1], tn[2], 0.f};
    n = mat_vec(f.cs_inv, n);
    n = mat_vec(ob.cs, n);
    uint32_t mark = 0;
    if (f.flags & J3DG_SHADOW) {
      // j3d/canvas.cpp:836-857 (any hit decides, only hit2.found is used)
      const uint32_t* tr = ob.idx + 3 * (size_t)best.tri;
      f4 V[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float* vv = ob.verts + 3 * (size_t)tr[k];
        V[k] = mat_vec(ob.cs, f4{vv[0], vv[1], vv[2], 1.f});
      }
      float s0 = 1.f - best.u - best.v;
      f4 pos = {V[0].x * s0 + best.u * V[1].x + best.v * V[2].x, V[0].y * s0 + best.u * V[1].y + best.v * V[2].y,
                V[0].z * s0 + best.u * V[1].z + best.v * V[2].z, V[0].w * s0 + best.u * V[1].w + best.v * V[2].w};
      f4 ld;
      if (f.flags & J3DG_DIRECTIONAL_LIGHT)
        ld = f4{f.light_dir[0], f.light_dir[1], f.light_dir[2], 0.f};
      else
        ld = f4{f.light[0] - pos.x, f.light[1] - pos.y, f.light[2] - pos.z, f.light[3] - pos.w};
      lstats[2]++;
      Hit sh;
      sh.found = false;
      for (uint32_t i = 0; i < f.n_obj; ++i) {
        const DevObject& o2 = objs[i];
        Ray r = make_object_ray(o2.inv_cs, pos, ld, 1e-3f, FLT_MAX);
        if (trace_object<W, D, true, COUNT>(o2.nodes, o2.tris, i, r, sh, st, c)) {
          mark |= 1u;
          break;
        }
      }
      if (mark) lstats[3]++;
    }
    *reinterpret_cast<uint8_t*>(pix) = (uint8_t)mark;
    float* pf = reinterpret_cast<float*>(pix);
    uint32_t* pu = reinterpret_cast<uint32_t*>(pix);
    pf[1] = n.x;
    pf[2] = n.y;
    pf[3] = best.t;
    pu[4] = best.tri;
    pf[5] = best.u;
    pf[6] = best.v;
    pu[7] = ob.db_id;
  } else {
    // j3d/canvas.cpp:859-866: mark, r, g, b and barycentrics are left as they were
    float* pf = reinterpret_cast<float*>(pix);
    uint32_t* pu = reinterpret_cast<uint32_t*>(pix);
    pf[1] = 0.f;
    pf[2] = 0.f;
    pf[3] = FLT_MAX;
    pu[4] = kMiss;
    pu[7] = 0u;
  }
}

__device__ inline void wave_add_stats(uint64_t* counters, const Counters& c, const uint32_t* lstats, bool COUNT) {
  // one atomic per wave and counter
  uint64_t v[6] = {lstats[0], lstats[1], lstats[2], lstats[3], COUNT ? c.nodes : 0, COUNT ? c.tris : 0};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    uint64_t s = v[k];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd((unsigned long long*)&counters[k], (unsigned long long)s);
  }
}

// Grid-stride (static) or atomic-queue (dynamic) persistent waves over 8x8 tiles.
template <int W, int D, bool COUNT, bool DYNAMIC>
__global__ void __launch_bounds__(256) k_trace(FrameParams f, const DevObject* __restrict__ objs, uint2* spill,
                                               uint32_t spill_cap, uint32_t* queue, uint64_t* counters) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave_in_block = threadIdx.x >> 6;
  const uint32_t waves_per_block = blockDim.x >> 6;
  const uint64_t total_lanes = (uint64_t)gridDim.x * blockDim.x;
  Stack<D> st;
  st.lds = lds_stack + threadIdx.x;
  st.stride = blockDim.x;
  st.spill = spill + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
  st.sstride = total_lanes;
  st.cap = spill_cap;
  st.sp = 0;
  st.overflow = false;
  Counters c = {0, 0};
  uint32_t lstats[4] = {0, 0, 0, 0};
  const uint32_t n_waves = gridDim.x * waves_per_block;
  uint32_t tile = blockIdx.x * waves_per_block + wave_in_block;
  if (DYNAMIC) {
    uint32_t t0 = 0;
    if (lane == 0) t0 = atomicAdd(queue, 1u);
    tile = __shfl(t0, 0, 64);
  }
  while (tile < f.n_tiles) {
    uint32_t px, py;
    if (tile_pixel(f, tile, lane, px, py)) trace_pixel<W, D, COUNT>(f, objs, px, py, st, c, lstats);
    if (DYNAMIC) {
      uint32_t t0 = 0;
      if (lane == 0) t0 = atomicAdd(queue, 1u);
      tile = __shfl(t0, 0, 64);
    } else {
      tile += n_waves;
    }
  }
  wave_add_stats(counters, c, lstats, COUNT);
  if (__any(st.overflow) && lane == 0) atomicAdd((unsigned long long*)&counters[6], 1ull);
}

__global__ void k_reset(uint32_t* queue, uint64_t* counters, uint32_t n_counters) {
  uint32_t i = threadIdx.x;
  if (i == 0) *queue = 0u;
  if (i < n_counters) counters[i] = 0ull;
}

// canvas::canvas_to_image (j3d/canvas.cpp:582-670) with the background copy of
// canvas::render_scene (canvas.cpp:895) fused in: one thread per pixel.
__global__ void __launch_bounds__(256) k_shade(ShadeParams s) {
  const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63);
  const uint32_t y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= s.w || y >= s.h) return;
  const j3dg_pixel* row = s.gbuf + (size_t)y * s.w;
  const j3dg_pixel p = row[x];
  uint32_t out;
  const uint32_t flags = s.flags;
  MatcapView m{s.matcap, s.mw, s.mh, s.cavity};
  if (p.object_id == kMiss) {
    out = background_color(y, s.h, s.bg_top, s.bg_bottom);
  } else if (flags & J3DG_ONE_BIT) {
    // j3d/canvas.cpp:498-579
    uint32_t clr = matcap_color(m, matcap_U(p.u, m.w), matcap_V(p.v, m.h), p.mark);
    int res = (int)((((clr & 0xff0000) >> 16) + ((clr & 0xff00) >> 8) + (clr & 0xff)) >> 7) + 1;
    if (x + 1 == s.w) {
      out = (((s.w - 1) % res == 0) && (y % res == 0)) ? 0xff000000u : 0xffffffffu;
    } else {
      const j3dg_pixel r = row[x + 1];
      const j3dg_pixel q = s.gbuf[(size_t)(y == 0 ? 0 : y - 1) * s.w + x];
      const float th = 0.001f;
      float angle = 1.f;
      bool er = r.object_id != kMiss && (fabsf(p.u - r.u) > th || fabsf(p.v - r.v) > th);
      bool eq = !er && q.object_id != kMiss && (fabsf(p.u - q.u) > th || fabsf(p.v - q.v) > th);
      if (er || eq) {
        float u2 = er ? r.u : q.u, v2 = er ? r.v : q.v;
        float w1 = sqrtf(1.f - p.u * p.u - p.v * p.v);
        float w2 = sqrtf(1.f - u2 * u2 - v2 * v2);
        angle = p.u * u2 + p.v * v2 + w1 * w2;
      }
      bool black = ((x % res == 0) && (y % res == 0));
      if (fabsf(angle) < 0.95f) black = (res == 1) ? !black : true;
      out = black ? 0xff000000u : 0xffffffffu;
    }
  } else if (flags & (J3DG_WIREFRAME | J3DG_EDGES)) {
    bool shading = (flags & J3DG_SHADING) != 0;
    if (x + 1 == s.w) {
      out = plain_color(p.mark, p.r, p.g, p.b, m, shading, p.u, p.v);
    } else {
      const j3dg_pixel r = row[x + 1];
      const j3dg_pixel q = s.gbuf[(size_t)(y == 0 ? 0 : y - 1) * s.w + x];
      if (flags & J3DG_WIREFRAME) {
        // j3d/canvas.cpp:449-496
        if ((r.object_id != kMiss && r.object_id != p.object_id) || (q.object_id != kMiss && q.object_id != p.object_id)) {
          const float scale = (p.u * p.u + p.v * p.v) * 0.5f;
          unsigned char cc = (unsigned char)(255.f * scale);
          out = make_color(cc, cc, cc);
        } else {
          out = plain_color(p.mark, p.r, p.g, p.b, m, shading, p.u, p.v);
        }
      } else {
        // j3d/canvas.cpp:590-649
        const float th = 0.001f;
        const float Wf = (float)s.w, Hf = (float)s.h;
        if (r.object_id != kMiss && (fabsf(p.u - r.u) > th || fabsf(p.v - r.v) > th)) {
          float angle = convex_cos_angle(s.proj_inv, s.near_plane, Wf, Hf, (float)x, (float)y, p.u, p.v, p.depth,
                                         (float)x + 1.f, (float)y, r.u, r.v, r.depth);
          out = angle_color(angle, m, p.u, p.v, p.mark);
        } else if (q.object_id != kMiss && (fabsf(p.u - q.u) > th || fabsf(p.v - q.v) > th)) {
          float angle = convex_cos_angle(s.proj_inv, s.near_plane, Wf, Hf, (float)x, (float)y, p.u, p.v, p.depth,
                                         (float)x, (float)y - 1.f, q.u, q.v, q.depth);
          out = angle_color(angle, m, p.u, p.v, p.mark);
        } else {
          out = plain_color(p.mark, p.r, p.g, p.b, m, shading, p.u, p.v);
        }
      }
    }
  } else {
    out = plain_color(p.mark, p.r, p.g, p.b, m, (flags & J3DG_SHADING) != 0, p.u, p.v);
  }
  s.rgba[(size_t)y * s.w + x] = out;
}

// Raw ray queries (world-space rays, closest or any hit over a list of objects).
template <int W, int D, bool ANY>
__global__ void __launch_bounds__(256) k_trace_rays(const float* __restrict__ rays, uint64_t n,
                                                    const DevObject* __restrict__ objs, uint32_t n_obj, uint2* spill,
                                                    uint32_t spill_cap, uint32_t* out_obj, uint32_t* out_tri,
                                                    float* out_t, float* out_u, float* out_v, uint8_t* out_any,
                                                    uint64_t* counters) {
  extern __shared__ uint2 lds_stack[];
  const uint64_t total_lanes = (uint64_t)gridDim.x * blockDim.x;
  Stack<D> st;
  st.lds = lds_stack + threadIdx.x;
  st.stride = blockDim.x;
  st.spill = spill + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
  st.sstride = total_lanes;
  st.cap = spill_cap;
  st.sp = 0;
  st.overflow = false;
  Counters c = {0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += total_lanes) {
    const float* rr = rays + 8 * i;
    f4 o = {rr[0], rr[1], rr[2], 1.f}, d = {rr[3], rr[4], rr[5], 0.f};
    Hit best;
    best.found = false;
    best.t = 0.f, best.u = 0.f, best.v = 0.f, best.tri = kMiss, best.obj = kMiss;
    bool any = false;
    for (uint32_t k = 0; k < n_obj; ++k) {
      Ray r = make_object_ray(objs[k].inv_cs, o, d, rr[6], rr[7]);
      if (trace_object<W, D, ANY, false>(objs[k].nodes, objs[k].tris, k, r, best, st, c)) {
        any = true;
        break;
      }
    }
    if (ANY) {
      out_any[i] = any ? 1 : 0;
    } else {
      out_obj[i] = best.found ? best.obj : kMiss;
      out_tri[i] = best.found ? best.tri : kMiss;
      out_t[i] = best.found ? best.t : FLT_MAX;
      out_u[i] = best.found ? best.u : 0.f;
      out_v[i] = best.found ? best.v : 0.f;
    }
  }
  if (__any(st.overflow) && (threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)&counters[6], 1ull);
}

__global__ void k_init_gbuf(j3dg_pixel* g, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  j3dg_pixel p;
  p.mark = p.r = p.g = p.b = 0;
  p.u = p.v = p.depth = 0.f;
  p.object_id = kMiss;
  p.barycentric_u = p.barycentric_v = 0.f;
  p.db_id = 0;
  g[i] = p;
}

// render_scene(nullptr) (j3d/canvas.cpp:883-890)
__global__ void k_no_scene(j3dg_pixel* g, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  g[i].db_id = 0;
  g[i].object_id = kMiss;
}

// ============================================================= host side ===

namespace {

#define HIP_TRY(expr, code)                                                         \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return ctx->fail(code, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <int W, int D>
struct TraceKernels {
  static void* get(bool count, bool dynamic) {
    if (count)
      return dynamic ? (void*)&k_trace<W, D, true, true> : (void*)&k_trace<W, D, true, false>;
    return dynamic ? (void*)&k_trace<W, D, false, true> : (void*)&k_trace<W, D, false, false>;
  }
};

void* trace_kernel(int width, int depth, bool count, bool dynamic) {
  if (width == 8) {
    if (depth == 8) return TraceKernels<8, 8>::get(count, dynamic);
    if (depth == 32) return TraceKernels<8, 32>::get(count, dynamic);
    return TraceKernels<8, 16>::get(count, dynamic);
  }
  if (depth == 8) return TraceKernels<4, 8>::get(count, dynamic);
  if (depth == 32) return TraceKernels<4, 32>::get(count, dynamic);
  return TraceKernels<4, 16>::get(count, dynamic);
}

}  // namespace

int Ctx::fail(int code, const std::string& msg) {
  last_error = msg;
  return code;
}

int Ctx::ensure_spill(uint64_t lanes, uint32_t entries) {
  uint64_t need = lanes * (uint64_t)entries;
  if (need <= spill_entries) return J3DG_OK;
  hipStreamSynchronize(stream);
  if (spill) hipFree(spill);
  spill = nullptr;
  spill_entries = 0;
  Ctx* ctx = this;
  HIP_TRY(hipMalloc(&spill, need * sizeof(uint2)), J3DG_ERR_OUT_OF_MEMORY);
  spill_entries = need;
  return J3DG_OK;
}

int Ctx::build_object_table(const uint32_t* slots, uint32_t n, uint32_t& max_stack, int& width) {
  std::vector<DevObject> tab(n);
  max_stack = 0;
  width = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (slots[i] >= meshes.size() || !meshes[slots[i]].live)
      return fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot " + std::to_string(slots[i]));
    const Mesh& m = meshes[slots[i]];
    if (width && width != (int)m.info.width) return fail(J3DG_ERR_UNSUPPORTED, "objects of one frame must share BVH width");
    width = (int)m.info.width;
    DevObject& o = tab[i];
    o.nodes = m.d_nodes;
    o.tris = m.d_tris;
    o.verts = m.d_verts;
    o.idx = m.d_idx;
    o.normals = m.d_normals;
    std::memcpy(o.cs, m.cs, sizeof(o.cs));
    std::memcpy(o.inv_cs, m.inv_cs, sizeof(o.inv_cs));
    o.db_id = m.db_id;
    o.width = m.info.width;
    max_stack = std::max(max_stack, m.max_stack);
  }
  if (n && tab.size() * sizeof(DevObject) == obj_cache.size() &&
      std::memcmp(obj_cache.data(), tab.data(), obj_cache.size()) == 0)
    return J3DG_OK;
  if (n > obj_capacity) {
    if (d_objs) hipFree(d_objs);
    d_objs = nullptr;
    obj_capacity = 0;
    Ctx* ctx = this;
    HIP_TRY(hipMalloc(&d_objs, std::max<uint32_t>(n, 16) * sizeof(DevObject)), J3DG_ERR_OUT_OF_MEMORY);
    obj_capacity = std::max<uint32_t>(n, 16);
  }
  if (n) {
    Ctx* ctx = this;
    HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
    HIP_TRY(hipMemcpy(d_objs, tab.data(), n * sizeof(DevObject), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  }
  obj_cache.assign(reinterpret_cast<const uint8_t*>(tab.data()),
                   reinterpret_cast<const uint8_t*>(tab.data()) + n * sizeof(DevObject));
  return J3DG_OK;
}

int Ctx::launch_trace(const FrameParams& fp, uint32_t max_stack, int width, bool learned) {
  (void)learned;
  Ctx* ctx = this;
  const uint32_t D = tuning.stack_depth;
  const uint32_t block = tuning.block_size;
  const uint32_t waves_per_block = block / 64;
  const uint32_t lds = D * block * sizeof(uint2);
  uint32_t blocks_needed = (fp.n_tiles + waves_per_block - 1) / waves_per_block;
  uint32_t per_cu = std::max(1u, std::min(160u * 1024u / std::max(lds, 1u), 32u / waves_per_block));
  uint32_t grid = std::max(1u, std::min(blocks_needed, (uint32_t)n_cus * per_cu));
  uint64_t lanes = (uint64_t)grid * block;
  uint32_t cap = max_stack > D ? max_stack - D + 1 : 1;
  int rc = ensure_spill(lanes, cap);
  if (rc) return rc;
  void* fn = trace_kernel(width, (int)D, counting, tuning.persistent != 0);
  uint2* sp = spill;
  uint32_t* q = d_queue;
  uint64_t* cnt = d_counters;
  FrameParams f = fp;
  const DevObject* objs = d_objs;
  void* args[] = {&f, &objs, &sp, &cap, &q, &cnt};
  HIP_TRY(hipEventRecord(ev[0], stream), J3DG_ERR_LAUNCH);
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, stream), J3DG_ERR_LAUNCH);
  HIP_TRY(hipEventRecord(ev[1], stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int Ctx::launch_shade(const ShadeParams& sp) {
  Ctx* ctx = this;
  dim3 grid((sp.w + 63) / 64, (sp.h + 3) / 4);
  HIP_TRY(hipEventRecord(ev[2], stream), J3DG_ERR_LAUNCH);
  hipLaunchKernelGGL(k_shade, grid, dim3(256), 0, stream, sp);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_TRY(hipEventRecord(ev[3], stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // namespace j3d

// ================================================================== C ABI ===

using j3d::Ctx;
using j3d::Mesh;
using j3d::Target;

#define CHECK_CTX()                         \
  if (!ctx) return J3DG_ERR_INVALID_ARG;    \
  std::lock_guard<std::mutex> lock_(ctx->mu)

#define HIP_CALL(expr, code)                                                        \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return ctx->fail(code, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" {

int j3dg_abi_version(void) { return J3DG_ABI_VERSION; }

int j3dg_device_count(int* count) {
  if (!count) return J3DG_ERR_INVALID_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return J3DG_OK;
}

int j3dg_create(int device, j3dg_ctx** out) {
  if (!out) return J3DG_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) return J3DG_ERR_NO_DEVICE;
  Ctx* ctx = new Ctx();
  ctx->device = device;
  auto cleanup = [&](int code) {
    j3dg_destroy(reinterpret_cast<j3dg_ctx*>(ctx));
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return cleanup(J3DG_ERR_NO_DEVICE);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return cleanup(J3DG_ERR_NO_DEVICE);
  ctx->n_cus = prop.multiProcessorCount;
  ctx->arch = prop.gcnArchName;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(J3DG_ERR_LAUNCH);
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(J3DG_ERR_LAUNCH);
  if (hipMalloc(&ctx->d_queue, 64) != hipSuccess) return cleanup(J3DG_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->d_counters, 16 * sizeof(uint64_t)) != hipSuccess) return cleanup(J3DG_ERR_OUT_OF_MEMORY);
  *out = reinterpret_cast<j3dg_ctx*>(ctx);
  return J3DG_OK;
}

void j3dg_destroy(j3dg_ctx* h) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto& m : ctx->meshes) m.release();
  for (auto& t : ctx->targets) t.release();
  ctx->release_buffers();
  for (auto& e : ctx->ev)
    if (e) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* j3dg_last_error(const j3dg_ctx* h) {
  const Ctx* ctx = reinterpret_cast<const Ctx*>(h);
  return ctx ? ctx->last_error.c_str() : "null context";
}

int j3dg_synchronize(j3dg_ctx* h) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_set_tuning(j3dg_ctx* h, const j3dg_tuning* t) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!t) return ctx->fail(J3DG_ERR_INVALID_ARG, "null tuning");
  j3dg_tuning n = *t;
  if (n.stack_depth == 0) n.stack_depth = 16;
  if (n.rays_per_item == 0) n.rays_per_item = 1;
  if (n.block_size == 0) n.block_size = 256;
  if (n.stack_depth != 8 && n.stack_depth != 16 && n.stack_depth != 32)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "stack_depth must be 8, 16 or 32");
  if (n.rays_per_item != 1 && n.rays_per_item != 2 && n.rays_per_item != 4)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "rays_per_item must be 1, 2 or 4");
  if (n.block_size != 64 && n.block_size != 128 && n.block_size != 256)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "block_size must be 64, 128 or 256");
  ctx->tuning = n;
  return J3DG_OK;
}

int j3dg_get_tuning(j3dg_ctx* h, j3dg_tuning* t) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!t) return ctx->fail(J3DG_ERR_INVALID_ARG, "null tuning");
  *t = ctx->tuning;
  return J3DG_OK;
}

int j3dg_set_counting(j3dg_ctx* h, int on) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  ctx->counting = on != 0;
  return J3DG_OK;
}

int j3dg_get_stats(j3dg_ctx* h, j3dg_stats* st) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!st) return ctx->fail(J3DG_ERR_INVALID_ARG, "null stats");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  uint64_t c[16];
  HIP_CALL(hipMemcpy(c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  std::memset(st, 0, sizeof(*st));
  st->primary_rays = c[0];
  st->hits = c[1];
  st->shadow_rays = c[2];
  st->shadowed = c[3];
  st->node_visits = c[4];
  st->tri_tests = c[5];
  st->stack_overflow = c[6];
  st->learned_resolved = c[7];
  st->learned_fallback = c[8];
  float ms = 0.f;
  if (ctx->traced && hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) st->trace_ms = ms;
  if (ctx->shaded && hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) st->shade_ms = ms;
  if (ctx->mlp_ran && hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]) == hipSuccess) st->mlp_ms = ms;
  return J3DG_OK;
}

int j3dg_upload_mesh(j3dg_ctx* h, uint32_t db_id, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                     uint64_t n_triangles, const float cs[16], const j3dg_build_opts* opts, uint32_t* slot) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!slot || (!xyz && n_vertices) || (!tris && n_triangles))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "null mesh array");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  BuildOptions bo;
  if (opts) {
    if (opts->width) bo.width = (int)opts->width;
    if (opts->leaf_size) bo.leaf_size = (int)opts->leaf_size;
    bo.threads = (int)opts->threads;
  }
  BuiltBVH bvh;
  const char* err = nullptr;
  if (!build_bvh(xyz, n_vertices, tris, n_triangles, bo, bvh, &err))
    return ctx->fail(J3DG_ERR_INVALID_ARG, err ? err : "BVH build failed");
  // per-triangle normals (j3d/scene.cpp:20) and bounding box (scene.cpp:22, mesh.cpp:39-58)
  std::vector<float> normals(3 * n_triangles);
  {
    int threads = bo.threads > 0 ? bo.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    uint64_t chunk = (n_triangles + threads - 1) / std::max(threads, 1);
    for (int t = 0; t < threads && chunk; ++t) {
      uint64_t a = std::min<uint64_t>(n_triangles, t * chunk), b = std::min<uint64_t>(n_triangles, (t + 1) * chunk);
      if (a < b)
        pool.emplace_back([&, a, b]() {
          for (uint64_t i = a; i < b; ++i)
            triangle_normal(xyz + 3 * (size_t)tris[3 * i], xyz + 3 * (size_t)tris[3 * i + 1],
                            xyz + 3 * (size_t)tris[3 * i + 2], &normals[3 * i]);
        });
    }
    for (auto& th : pool) th.join();
  }
  Mesh m;
  m.live = true;
  m.db_id = db_id;
  if (cs)
    std::memcpy(m.cs, cs, sizeof(m.cs));
  else
    for (int i = 0; i < 16; ++i) m.cs[i] = (i % 5 == 0) ? 1.f : 0.f;
  invert_orthonormal(m.cs, m.inv_cs);
  m.info.n_vertices = n_vertices;
  m.info.n_triangles = n_triangles;
  m.info.n_nodes = bvh.n_nodes;
  m.info.width = (uint32_t)bvh.width;
  m.info.max_depth = bvh.max_depth;
  m.max_stack = bvh.max_stack + 1;
  if (n_vertices) {
    for (int k = 0; k < 3; ++k) m.info.bb_min[k] = m.info.bb_max[k] = xyz[k];
    for (uint64_t i = 1; i < n_vertices; ++i)
      for (int k = 0; k < 3; ++k) {
        m.info.bb_min[k] = std::min(m.info.bb_min[k], xyz[3 * i + k]);
        m.info.bb_max[k] = std::max(m.info.bb_max[k], xyz[3 * i + k]);
      }
  }
  auto t1 = std::chrono::steady_clock::now();
  m.info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
  m.normals_host = std::move(normals);
  size_t nb = bvh.nodes.size() * sizeof(float), tb = bvh.tris.size() * sizeof(float);
  size_t vb = 3 * n_vertices * sizeof(float), ib = 3 * n_triangles * sizeof(uint32_t);
  size_t mb = 3 * n_triangles * sizeof(float);
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)); };
  if (alloc((void**)&m.d_nodes, nb) != hipSuccess || alloc((void**)&m.d_tris, tb) != hipSuccess ||
      alloc((void**)&m.d_verts, vb) != hipSuccess || alloc((void**)&m.d_idx, ib) != hipSuccess ||
      alloc((void**)&m.d_normals, mb) != hipSuccess) {
    m.release();
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device allocation for mesh failed");
  }
  bool ok = hipMemcpy(m.d_nodes, bvh.nodes.data(), nb, hipMemcpyHostToDevice) == hipSuccess &&
            (!tb || hipMemcpy(m.d_tris, bvh.tris.data(), tb, hipMemcpyHostToDevice) == hipSuccess) &&
            (!vb || hipMemcpy(m.d_verts, xyz, vb, hipMemcpyHostToDevice) == hipSuccess) &&
            (!ib || hipMemcpy(m.d_idx, tris, ib, hipMemcpyHostToDevice) == hipSuccess) &&
            (!mb || hipMemcpy(m.d_normals, m.normals_host.data(), mb, hipMemcpyHostToDevice) == hipSuccess);
  if (!ok) {
    m.release();
    return ctx->fail(J3DG_ERR_LAUNCH, "mesh upload copy failed");
  }
  auto t2 = std::chrono::steady_clock::now();
  m.info.upload_seconds = std::chrono::duration<double>(t2 - t1).count();
  m.info.device_bytes = nb + tb + vb + ib + mb;
  uint32_t s = (uint32_t)ctx->meshes.size();
  for (uint32_t i = 0; i < ctx->meshes.size(); ++i)
    if (!ctx->meshes[i].live) {
      s = i;
      break;
    }
  if (s == ctx->meshes.size()) ctx->meshes.emplace_back();
  ctx->meshes[s] = std::move(m);
  ctx->obj_cache.clear();
  *slot = s;
  return J3DG_OK;
}

int j3dg_remove_mesh(j3dg_ctx* h, uint32_t slot) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  ctx->meshes[slot].release();
  ctx->obj_cache.clear();
  ctx->invalidate_graphs();
  return J3DG_OK;
}

int j3dg_set_mesh_cs(j3dg_ctx* h, uint32_t slot, const float cs[16]) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!cs) return ctx->fail(J3DG_ERR_INVALID_ARG, "null matrix");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  std::memcpy(ctx->meshes[slot].cs, cs, 16 * sizeof(float));
  invert_orthonormal(ctx->meshes[slot].cs, ctx->meshes[slot].inv_cs);
  return J3DG_OK;
}

int j3dg_get_mesh_info(j3dg_ctx* h, uint32_t slot, j3dg_mesh_info* info) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!info) return ctx->fail(J3DG_ERR_INVALID_ARG, "null info");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  *info = ctx->meshes[slot].info;
  return J3DG_OK;
}

int j3dg_mesh_normals(j3dg_ctx* h, uint32_t slot, float* out) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!out) return ctx->fail(J3DG_ERR_INVALID_ARG, "null output");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  const auto& n = ctx->meshes[slot].normals_host;
  if (!n.empty()) std::memcpy(out, n.data(), n.size() * sizeof(float));
  return J3DG_OK;
}

int j3dg_set_matcap(j3dg_ctx* h, const uint32_t* im, uint32_t w, uint32_t hgt, uint32_t cavity) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!im || w == 0 || hgt == 0 || (uint64_t)w * hgt > (1ull << 26)) return ctx->fail(J3DG_ERR_INVALID_ARG, "bad matcap");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  if ((uint64_t)w * hgt > ctx->matcap_capacity) {
    if (ctx->d_matcap) hipFree(ctx->d_matcap);
    ctx->d_matcap = nullptr;
    ctx->matcap_capacity = 0;
    HIP_CALL(hipMalloc(&ctx->d_matcap, (size_t)w * hgt * sizeof(uint32_t)), J3DG_ERR_OUT_OF_MEMORY);
    ctx->matcap_capacity = (uint64_t)w * hgt;
  }
  HIP_CALL(hipMemcpy(ctx->d_matcap, im, (size_t)w * hgt * sizeof(uint32_t), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  ctx->matcap_w = w;
  ctx->matcap_h = hgt;
  ctx->cavity = cavity;
  ctx->invalidate_graphs();
  return J3DG_OK;
}

int j3dg_create_target(j3dg_ctx* h, uint32_t w, uint32_t hgt, uint32_t* target) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!target || w == 0 || hgt == 0 || (uint64_t)w * hgt > (1ull << 31)) return ctx->fail(J3DG_ERR_INVALID_ARG, "bad target size");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  Target t;
  t.live = true;
  t.w = w;
  t.h = hgt;
  uint64_t n = (uint64_t)w * hgt;
  if (hipMalloc(&t.gbuf, n * sizeof(j3dg_pixel)) != hipSuccess || hipMalloc(&t.rgba, n * sizeof(uint32_t)) != hipSuccess) {
    t.release();
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device allocation for target failed");
  }
  hipLaunchKernelGGL(k_init_gbuf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, t.gbuf, n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemsetAsync(t.rgba, 0, n * sizeof(uint32_t), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    t.release();
    return ctx->fail(J3DG_ERR_LAUNCH, std::string("target init: ") + hipGetErrorString(e));
  }
  uint32_t s = (uint32_t)ctx->targets.size();
  for (uint32_t i = 0; i < ctx->targets.size(); ++i)
    if (!ctx->targets[i].live) {
      s = i;
      break;
    }
  if (s == ctx->targets.size()) ctx->targets.emplace_back();
  ctx->targets[s] = t;
  *target = s;
  return J3DG_OK;
}

int j3dg_destroy_target(j3dg_ctx* h, uint32_t target) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  ctx->targets[target].release();
  ctx->invalidate_graphs();
  return J3DG_OK;
}

int j3dg_write_target_gbuffer(j3dg_ctx* h, uint32_t target, const j3dg_pixel* gbuf) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!gbuf) return ctx->fail(J3DG_ERR_INVALID_ARG, "null gbuffer");
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& t = ctx->targets[target];
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(t.gbuf, gbuf, (size_t)t.w * t.h * sizeof(j3dg_pixel), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_target_device_ptrs(j3dg_ctx* h, uint32_t target, void** gbuf, void** rgba) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  if (gbuf) *gbuf = ctx->targets[target].gbuf;
  if (rgba) *rgba = ctx->targets[target].rgba;
  return J3DG_OK;
}

int j3dg_read_target(j3dg_ctx* h, uint32_t target, j3dg_pixel* gbuf, uint32_t* rgba) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& t = ctx->targets[target];
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  size_t n = (size_t)t.w * t.h;
  if (gbuf) HIP_CALL(hipMemcpy(gbuf, t.gbuf, n * sizeof(j3dg_pixel), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  if (rgba) HIP_CALL(hipMemcpy(rgba, t.rgba, n * sizeof(uint32_t), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_render(j3dg_ctx* h, const j3dg_frame_desc* fd, uint32_t target, uint32_t stages, j3dg_pixel* gbuf_out,
                uint32_t* rgba_out) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!fd) return ctx->fail(J3DG_ERR_INVALID_ARG, "null frame");
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& tg = ctx->targets[target];
  if (fd->width != tg.w || fd->height != tg.h) return ctx->fail(J3DG_ERR_INVALID_ARG, "frame size differs from target size");
  if (fd->n_objects > J3DG_MAX_OBJECTS) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many objects");
  if ((stages & J3DG_STAGE_SHADE) && !ctx->d_matcap) return ctx->fail(J3DG_ERR_INVALID_ARG, "no matcap set");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  ctx->traced = ctx->shaded = ctx->mlp_ran = false;
  const uint64_t npx = (uint64_t)tg.w * tg.h;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->d_queue, ctx->d_counters, 16u);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  if (stages & J3DG_STAGE_TRACE) {
    if (fd->flags & J3DG_NO_SCENE) {
      hipLaunchKernelGGL(k_no_scene, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, ctx->stream, tg.gbuf, npx);
      HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
    } else {
      FrameParams fp;
      int rc = make_frame_params(*fd, tg, fp);
      if (rc) return ctx->fail(rc, "invalid frame");
      uint32_t max_stack = 0;
      int width = 4;
      if (fd->n_objects == 0) {
        // j3d/canvas.cpp:744-760: an empty scene writes all-miss pixels
        fp.n_obj = 0;
        width = 4;
      } else {
        rc = ctx->build_object_table(fd->objects, fd->n_objects, max_stack, width);
        if (rc) return rc;
      }
      rc = ctx->launch_trace(fp, std::max<uint32_t>(max_stack, 1), width, (fd->flags & J3DG_LEARNED) != 0);
      if (rc) return rc;
      ctx->traced = true;
    }
  }
  if (stages & J3DG_STAGE_SHADE) {
    ShadeParams sp;
    make_shade_params(*fd, tg, *ctx, sp);
    int rc = ctx->launch_shade(sp);
    if (rc) return rc;
    ctx->shaded = true;
  }
  if (gbuf_out) HIP_CALL(hipMemcpyAsync(gbuf_out, tg.gbuf, npx * sizeof(j3dg_pixel), hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
  if (rgba_out) HIP_CALL(hipMemcpyAsync(rgba_out, tg.rgba, npx * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
  if (gbuf_out || rgba_out) HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

static int trace_rays_impl(Ctx* ctx, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                           uint32_t* obj, uint32_t* tri, float* t, float* u, float* v, uint8_t* hit, bool any) {
  if (!rays && n) return ctx->fail(J3DG_ERR_INVALID_ARG, "null rays");
  if (n_slots > J3DG_MAX_OBJECTS || (n_slots && !slots)) return ctx->fail(J3DG_ERR_INVALID_ARG, "bad object list");
  if (n == 0) return J3DG_OK;
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  uint32_t max_stack = 1;
  int width = 4;
  if (n_slots) {
    int rc = ctx->build_object_table(slots, n_slots, max_stack, width);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->d_queue, ctx->d_counters, 16u);
  float* d_rays = nullptr;
  uint8_t* d_out = nullptr;
  size_t out_bytes = any ? n : n * 20;
  HIP_CALL(hipMalloc(&d_rays, n * 8 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&d_out, out_bytes) != hipSuccess) {
    hipFree(d_rays);
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "ray output allocation failed");
  }
  struct Guard {
    void* a;
    void* b;
    ~Guard() {
      hipFree(a);
      hipFree(b);
    }
  } guard{d_rays, d_out};
  HIP_CALL(hipMemcpyAsync(d_rays, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice, ctx->stream), J3DG_ERR_LAUNCH);
  const uint32_t D = 16, block = 256;
  uint64_t blocks = (n + block - 1) / block;
  uint32_t grid = (uint32_t)std::min<uint64_t>(blocks, (uint64_t)ctx->n_cus * 4);
  uint32_t cap = max_stack > D ? max_stack - D + 1 : 1;
  int rc = ctx->ensure_spill((uint64_t)grid * block, cap);
  if (rc) return rc;
  uint32_t* o_obj = reinterpret_cast<uint32_t*>(d_out);
  uint32_t* o_tri = o_obj + n;
  float* o_t = reinterpret_cast<float*>(o_tri + n);
  float* o_u = o_t + n;
  float* o_v = o_u + n;
  uint8_t* o_any = d_out;
  const DevObject* objs = ctx->d_objs;
  uint32_t no = n_slots;
  size_t lds = D * block * sizeof(uint2);
  if (width == 8) {
    if (any)
      hipLaunchKernelGGL((k_trace_rays<8, 16, true>), dim3(grid), dim3(block), lds, ctx->stream, d_rays, n, objs, no, ctx->spill, cap, o_obj, o_tri, o_t, o_u, o_v, o_any, ctx->d_counters);
    else
      hipLaunchKernelGGL((k_trace_rays<8, 16, false>), dim3(grid), dim3(block), lds, ctx->stream, d_rays, n, objs, no, ctx->spill, cap, o_obj, o_tri, o_t, o_u, o_v, o_any, ctx->d_counters);
  } else {
    if (any)
      hipLaunchKernelGGL((k_trace_rays<4, 16, true>), dim3(grid), dim3(block), lds, ctx->stream, d_rays, n, objs, no, ctx->spill, cap, o_obj, o_tri, o_t, o_u, o_v, o_any, ctx->d_counters);
    else
      hipLaunchKernelGGL((k_trace_rays<4, 16, false>), dim3(grid), dim3(block), lds, ctx->stream, d_rays, n, objs, no, ctx->spill, cap, o_obj, o_tri, o_t, o_u, o_v, o_any, ctx->d_counters);
  }
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  if (any) {
    HIP_CALL(hipMemcpy(hit, o_any, n, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  } else {
    if (obj) HIP_CALL(hipMemcpy(obj, o_obj, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (tri) HIP_CALL(hipMemcpy(tri, o_tri, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (t) HIP_CALL(hipMemcpy(t, o_t, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (u) HIP_CALL(hipMemcpy(u, o_u, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (v) HIP_CALL(hipMemcpy(v, o_v, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  }
  uint64_t c[16];
  HIP_CALL(hipMemcpy(c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  if (c[6]) return ctx->fail(J3DG_ERR_INTERNAL, "traversal stack overflow");
  return J3DG_OK;
}

int j3dg_trace_rays(j3dg_ctx* h, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                    uint32_t* obj, uint32_t* tri, float* t, float* u, float* v) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  return trace_rays_impl(ctx, slots, n_slots, rays, n, obj, tri, t, u, v, nullptr, false);
}

int j3dg_any_hit_rays(j3dg_ctx* h, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                      uint8_t* hit) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!hit && n) return ctx->fail(J3DG_ERR_INVALID_ARG, "null output");
  return trace_rays_impl(ctx, slots, n_slots, rays, n, nullptr, nullptr, nullptr, nullptr, nullptr, hit, true);
}

int j3dg_render_batch(j3dg_ctx* h, const j3dg_frame_desc* frames, const uint32_t* targets, uint32_t n) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  (void)frames;
  (void)targets;
  (void)n;
  return ctx->fail(J3DG_ERR_UNSUPPORTED, "batch rendering not built yet");
}

}  // extern "C"
