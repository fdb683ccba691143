// traverse.h -- device-side BVH traversal for gfx950.
//
// One ray per lane of a wave64.  The traversal is a per-lane state machine
// advanced one step (one wide node or one leaf) per loop trip, so that a
// persistent wave can hand a finished lane a new ray at once ("refill")
// instead of idling it until the slowest lane of the wave is done.
//
// Closest hit is the lexicographic minimum of (t, object index, triangle id)
// over every triangle the Moller-Trumbore test accepts, so the result does
// not depend on the BVH topology or the visiting order.  Box culling is
// conservative: child boxes are widened by 2^-20 of their magnitude at build
// time and every t comparison carries a 2^-16 relative slack.
#pragma once
#include <hip/hip_runtime.h>

#include "bvh_build.h"
#include "j3dg_internal.h"
#include "rt_math.h"

namespace j3d {

constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kSlackUp = 1.0f + 1.0f / 65536.0f;
constexpr float kSlackDown = 1.0f - 1.0f / 65536.0f;

struct Hit {
  float t, u, v;
  uint32_t tri, obj;
  bool found;
};

struct Ray {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz;  // reciprocal direction, zero components replaced by +-1e-30
  float t_near, t_far;
};

// Loads through the global address space.  Pointers read from the object table
// are generic to the compiler, which would otherwise emit flat loads (they
// wait on the LDS counter too, serialising them against the LDS stack).
template <class T>
__device__ inline T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// float4 is a class type whose copy goes through a generic reference, so the
// 16-byte loads use the native vector type.
typedef float j3d_f32x4 __attribute__((ext_vector_type(4)));
__device__ inline float4 gload4(const float4* p) {
  j3d_f32x4 v = *(const __attribute__((address_space(1))) j3d_f32x4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// Reciprocal of a direction component for the slab tests (|d| clamped to
// 1e-30 so no slab distance is inf * 0).  Only culling uses it, and culling
// widens every box by a relative 2^-16 (kSlackUp/kSlackDown), so the hardware
// reciprocal's 1-ulp error (2^-23 relative, scaling a node's entry and exit
// distances alike) cannot drop a box the ray enters; no hit depends on it.
__device__ inline float safe_rcp(float d) {
  float a = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
  return __builtin_amdgcn_rcpf(a);
}

// Primary rays come in tiles of 8x8 pixels: ray id = tile * 64 + lane.
__device__ inline bool tile_pixel(const FrameParams& f, uint32_t tile, uint32_t l, uint32_t& px, uint32_t& py) {
  uint32_t tx = tile % f.tiles_x, ty = tile / f.tiles_x;
  px = (uint32_t)f.x0 + tx * 8 + (l & 7);
  py = (uint32_t)f.y0 + ty * 8 + (l >> 3);
  return px <= (uint32_t)f.x1 && py <= (uint32_t)f.y1;
}

// World-space primary ray of pixel (px, py) (j3d/canvas.cpp:773-782): the
// screen point through P^-1 (camera direction with z = -1), then CS.
__device__ inline void primary_ray(const FrameParams& f, uint32_t px, uint32_t py, f4& wo, f4& wd) {
  const float Wf = (float)f.w, Hf = (float)f.h;
  f4 sp = {2.f * (((float)px + 0.5f) / Wf) - 1.f, 2.f * (((float)py + 0.5f) / Hf) - 1.f, f.near_plane, 1.f};
  f4 dir = mat_vec(f.proj_inv, sp);
  dir.w = 0.f;
  wd = mat_vec(f.cs, dir);
  wo = mat_vec(f.cs, f4{0.f, 0.f, 0.f, 1.f});
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

// Per-lane traversal stack: the top D entries live in LDS (interleaved by lane,
// so any mix of stack pointers is bank-conflict free); older entries spill to a
// global scratch area laid out [entry][lane] so that one wave's spills coalesce.
// Entries are 2 x 32 bits (child reference, entry distance), accessed through
// explicitly LDS- and global-typed pointers: a generic pointer would compile to
// flat instructions, which wait on the vector memory counter as well and so
// serialise every push and pop against the lane's outstanding loads and stores.
typedef uint32_t j3d_u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) j3d_u32x2 lds_entry;
typedef __attribute__((address_space(1))) j3d_u32x2 glb_entry;

template <int D>
struct Stack {
  lds_entry* lds;    // &lds[0][tid], stride `stride`
  uint32_t stride;   // threads per block
  glb_entry* spill;  // &spill[0][global lane]
  uint64_t sstride;  // lanes of the grid
  uint32_t cap;      // spill entries per lane
  uint32_t sp;
  bool overflow;

  __device__ void init(uint2* lds_base, uint2* spill_base, uint32_t spill_cap) {
    init(lds_base, spill_base, spill_cap, threadIdx.x);
  }
  // an empty stack in the LDS and spill column of thread `col` of the block
  __device__ void init(uint2* lds_base, uint2* spill_base, uint32_t spill_cap, uint32_t col) {
    lds = (lds_entry*)lds_base + col;
    stride = blockDim.x;
    sstride = (uint64_t)gridDim.x * blockDim.x;
    spill = (glb_entry*)spill_base + ((uint64_t)blockIdx.x * blockDim.x + col);
    cap = spill_cap;
    sp = 0;
    overflow = false;
  }
  // stores entry e (0-based from the bottom) of this lane; false on overflow
  __device__ bool put(uint32_t e, uint32_t ref, float dist) {
    const uint32_t slot = e & (D - 1);
    if (e >= (uint32_t)D) {
      if (e - D >= cap) {  // cap comes from the builder's stack bound; never expected
        overflow = true;
        return false;
      }
      spill[(uint64_t)(e - D) * sstride] = lds[slot * stride];
    }
    // lanes sharing one stack (the cooperative tail) all read the slot for
    // the spill before any of them overwrites it; see pop
    __builtin_amdgcn_wave_barrier();
    j3d_u32x2 v = {ref, __float_as_uint(dist)};
    lds[slot * stride] = v;
    return true;
  }
  __device__ void push(uint32_t ref, float dist) {
    if (put(sp, ref, dist)) ++sp;
  }
  __device__ bool pop(uint32_t& ref, float& dist) {
    if (sp == 0) return false;
    --sp;
    const uint32_t slot = sp & (D - 1);
    const j3d_u32x2 v = lds[slot * stride];
    // In the cooperative tail the lanes of a group pop one shared stack: each
    // reads the top entry before any of them writes the spilled entry below it
    // back into that slot.  The wave executes both in lockstep; the barrier
    // (no instruction) keeps the compiler from moving the write above the read.
    __builtin_amdgcn_wave_barrier();
    if (sp >= (uint32_t)D) lds[slot * stride] = spill[(uint64_t)(sp - D) * sstride];
    ref = v.x;
    dist = __uint_as_float(v.y);
    return true;
  }
};

struct Counters {
  uint32_t nodes, tris, steps;
};

// Visits one wide node: pushes the entered children far-to-near and returns the
// nearest (kNone if no child is entered).
__device__ inline float byte_f(uint32_t w, int i) { return (float)((w >> (8 * i)) & 0xFFu); }

// Slab tests of one axis for 4 children packed one byte each in (near, far)
// quantized words: t = q * scale + bias with scale = 2^e / d and
// bias = (p - o) / d, one fused multiply-add per plane.
__device__ inline void axis4(uint32_t qn, uint32_t qf, float scale, float bias, float* tn, float* tf) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    tn[i] = fmaf(byte_f(qn, i), scale, bias);
    tf[i] = fmaf(byte_f(qf, i), scale, bias);
  }
}

// Slab test of children [4h, 4h+4) of a quantized wide node (layout in bvh_build.h).
__device__ inline void test4(uint32_t lox, uint32_t hix, uint32_t loy, uint32_t hiy, uint32_t loz, uint32_t hiz,
                             const float* sc, const float* bs, const Ray& r, float tlo, float thi, float* dist) {
  float tnx[4], tfx[4], tny[4], tfy[4], tnz[4], tfz[4];
  // the near plane is the low one for a positive direction component
  if (r.ix < 0.f) axis4(hix, lox, sc[0], bs[0], tnx, tfx); else axis4(lox, hix, sc[0], bs[0], tnx, tfx);
  if (r.iy < 0.f) axis4(hiy, loy, sc[1], bs[1], tny, tfy); else axis4(loy, hiy, sc[1], bs[1], tny, tfy);
  if (r.iz < 0.f) axis4(hiz, loz, sc[2], bs[2], tnz, tfz); else axis4(loz, hiz, sc[2], bs[2], tnz, tfz);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float en = fmaxf(fmaxf(tnx[i], tny[i]), fmaxf(tnz[i], tlo));
    float ex = fminf(fminf(fminf(tfx[i], tfy[i]), tfz[i]) * kSlackUp, thi);
    dist[i] = en <= ex ? en : INFINITY;
  }
}

// Visits one wide node: pushes the entered children far-to-near and returns the
// nearest (kNone if no child is entered).
template <int W, int D, bool COUNT>
__device__ inline uint32_t decode_node(const float4 (&v)[W == 8 ? 6 : 4], const Ray& r, float thi, Stack<D>& st,
                                       Counters& c, bool ordered = true);

template <int W, int D, bool COUNT>
__device__ inline uint32_t visit_node(const float4* __restrict__ nodes, uint32_t node, const Ray& r, float thi,
                                      Stack<D>& st, Counters& c) {
  constexpr int kVec = W == 8 ? 6 : 4;  // 16-byte loads per node
  const float4* np = nodes + (size_t)node * kVec;
  float4 v[kVec];
#pragma unroll
  for (int k = 0; k < kVec; ++k) v[k] = gload4(np + k);
  return decode_node<W, D, COUNT>(v, r, thi, st, c);
}

// Tests the node's children whose 16-byte words are in v, pushes the entered
// ones far-to-near and returns the nearest (kNone if none is entered).
template <int W, int D, bool COUNT>
__device__ inline uint32_t decode_node(const float4 (&v)[W == 8 ? 6 : 4], const Ray& r, float thi, Stack<D>& st,
                                       Counters& c, bool ordered) {
  if (COUNT) c.nodes++;
  const float tlo = r.t_near * kSlackDown;
  const uint32_t ex = __float_as_uint(v[0].w);
  float sc[3], bs[3];
  sc[0] = __uint_as_float((ex & 0xFFu) << 23) * r.ix;
  sc[1] = __uint_as_float(((ex >> 8) & 0xFFu) << 23) * r.iy;
  sc[2] = __uint_as_float(((ex >> 16) & 0xFFu) << 23) * r.iz;
  bs[0] = (v[0].x - r.ox) * r.ix;
  bs[1] = (v[0].y - r.oy) * r.iy;
  bs[2] = (v[0].z - r.oz) * r.iz;
  float dist[W];
  uint32_t ref[W];
#define U(x) __float_as_uint(x)
  if (W == 4) {
    test4(U(v[1].x), U(v[1].y), U(v[1].z), U(v[1].w), U(v[2].x), U(v[2].y), sc, bs, r, tlo, thi, dist);
    ref[0] = U(v[2].z), ref[1] = U(v[2].w), ref[2] = U(v[3].x), ref[3] = U(v[3].y);
  } else {
    // words: x = v[1] (lo0, lo1, hi0, hi1), y = v[2], z = v[3]
    test4(U(v[1].x), U(v[1].z), U(v[2].x), U(v[2].z), U(v[3].x), U(v[3].z), sc, bs, r, tlo, thi, dist);
    test4(U(v[1].y), U(v[1].w), U(v[2].y), U(v[2].w), U(v[3].y), U(v[3].w), sc, bs, r, tlo, thi, dist + 4);
    ref[0] = U(v[4].x), ref[1] = U(v[4].y), ref[2] = U(v[4].z), ref[3] = U(v[4].w);
    ref[4] = U(v[5].x), ref[5] = U(v[5].y), ref[6] = U(v[5].z), ref[7] = U(v[5].w);
  }
#undef U
#define J3D_CSWAP(a, bb)           \
  if (dist[bb] < dist[a]) {         \
    float td = dist[a];             \
    dist[a] = dist[bb];             \
    dist[bb] = td;                  \
    uint32_t tr = ref[a];           \
    ref[a] = ref[bb];               \
    ref[bb] = tr;                   \
  }
  // Any-hit rays (shadows) need no order: when no lane of the wave wants the
  // nearest child first, the sorting network is skipped and the entered
  // children are pushed in slot order.
  if (W == 4) {
    J3D_CSWAP(0, 1) J3D_CSWAP(2, 3) J3D_CSWAP(0, 2) J3D_CSWAP(1, 3) J3D_CSWAP(1, 2)
  } else if (__ballot(ordered)) {
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

// Whether the ray enters any child box of the wide node at np within
// [t_near, thi] -- the first test a traversal makes at the root: a ray that
// enters none of the root's children hits nothing in the object.
template <int W>
__device__ inline bool enters_node(const float4* __restrict__ np, const Ray& r, float thi) {
  constexpr int kVec = W == 8 ? 6 : 4;
  float4 v[kVec];
#pragma unroll
  for (int k = 0; k < kVec; ++k) v[k] = gload4(np + k);
  const float tlo = r.t_near * kSlackDown;
  const uint32_t ex = __float_as_uint(v[0].w);
  float sc[3], bs[3];
  sc[0] = __uint_as_float((ex & 0xFFu) << 23) * r.ix;
  sc[1] = __uint_as_float(((ex >> 8) & 0xFFu) << 23) * r.iy;
  sc[2] = __uint_as_float(((ex >> 16) & 0xFFu) << 23) * r.iz;
  bs[0] = (v[0].x - r.ox) * r.ix;
  bs[1] = (v[0].y - r.oy) * r.iy;
  bs[2] = (v[0].z - r.oz) * r.iz;
  float dist[W];
#define U(x) __float_as_uint(x)
  if constexpr (W == 4) {
    test4(U(v[1].x), U(v[1].y), U(v[1].z), U(v[1].w), U(v[2].x), U(v[2].y), sc, bs, r, tlo, thi, dist);
  } else {
    test4(U(v[1].x), U(v[1].z), U(v[2].x), U(v[2].z), U(v[3].x), U(v[3].z), sc, bs, r, tlo, thi, dist);
    test4(U(v[1].y), U(v[1].w), U(v[2].y), U(v[2].w), U(v[3].y), U(v[3].w), sc, bs, r, tlo, thi, dist + 4);
  }
#undef U
  bool any = false;
#pragma unroll
  for (int i = 0; i < W; ++i) any |= dist[i] != INFINITY;
  return any;
}

// Traversal state of one lane inside one object.
struct Trav {
  Ray r;
  const float4* nodes;
  const float4* tris;
  float tmax;   // acceptance bound: min(t_far, best t)
  float thi;    // culling bound: tmax * (1 + 2^-16)
  uint32_t cur; // next node / leaf reference, kNone when the stack must be popped
  uint32_t obj; // object index
  uint32_t base;  // scenes with a top level: stack entries below this one are the top level's
  bool any;     // any-hit (shadow) mode
};

// A ray whose origin or direction is not finite hits nothing (the triangle
// test rejects it), and it must not be traversed either: its slab tests give
// NaN, which fminf / fmaxf drop, so every child -- empty slots, whose ref is
// kNone, included -- would count as entered.  begin_object gives such a ray a
// finite stand-in with an empty t range (t_near = +inf): the root's children
// all test as not entered, so it visits the root only and finishes as a miss.
__device__ inline bool ray_finite(f4 o, f4 d) {
  const float m = fmaxf(fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), fabsf(d.x))),
                        fmaxf(fabsf(d.y), fabsf(d.z)));
  return m < INFINITY && o.x == o.x && o.y == o.y && o.z == o.z && d.x == d.x && d.y == d.y && d.z == d.z;
}

// Starts traversal of object `oi` for the world ray (o, d) over [tn, tf].
__device__ inline void begin_object(Trav& s, const DevObject* __restrict__ objs, uint32_t oi, f4 o, f4 d, float tn,
                                    float tf, const Hit& best) {
  const DevObject& ob = objs[oi];
  if (!ray_finite(o, d)) {
    o = f4{0.f, 0.f, 0.f, 1.f};
    d = f4{1.f, 0.f, 0.f, 0.f};
    tn = INFINITY;
  }
  s.r = make_object_ray(ob.inv_cs, o, d, tn, tf);
  s.nodes = ob.nodes;
  s.tris = ob.tris;
  s.obj = oi;
  s.tmax = best.found ? fminf(best.t, tf) : tf;
  s.thi = s.tmax * kSlackUp;
  s.cur = 0;  // the root is always an internal node
}

// One traversal step: visits s.cur (a wide node or a leaf), then advances
// s.cur, popping (with culling against the current bound) if needed.
// Returns true when this object's traversal is complete (or, in any-hit mode,
// as soon as a hit is found).
// Tests the triangle record (a = v0 | id, e1, e2) against the lane's ray and
// keeps the closest hit.  Returns true when an any-hit ray is done.
template <bool COUNT>
__device__ inline bool test_triangle(Trav& s, Hit& best, Counters& c, const float4& a, const float4& e1,
                                     const float4& e2) {
  if (COUNT) c.tris++;
  float t, u, v;
  if (intersect_triangle(s.r.ox, s.r.oy, s.r.oz, s.r.dx, s.r.dy, s.r.dz, a.x, a.y, a.z, e1.x, e1.y, e1.z, e2.x,
                         e2.y, e2.z, s.r.t_near, s.tmax, t, u, v)) {
    uint32_t id = __float_as_uint(a.w);
    if (s.any) {
      best.found = true;
      best.t = t, best.u = u, best.v = v, best.tri = id, best.obj = s.obj;
      return true;
    }
    bool better = !best.found || t < best.t ||
                  (t == best.t && (s.obj < best.obj || (s.obj == best.obj && id < best.tri)));
    if (better) {
      best.found = true;
      best.t = t, best.u = u, best.v = v, best.tri = id, best.obj = s.obj;
      s.tmax = t;
      s.thi = t * kSlackUp;
    }
  }
  return false;
}

// Pops entries (culled against the current bound) until one is entered.
// Returns true when the stack runs dry: this object's traversal is complete.
template <int D, bool TOP = false>
__device__ inline bool pop_next(Trav& s, Stack<D>& st) {
  uint32_t ref;
  float dist;
  while ((!TOP || st.sp > s.base) && st.pop(ref, dist)) {
    if (dist <= s.thi) {
      s.cur = ref;
      return false;
    }
  }
  return true;
}

// 8-wide trees: an 8-wide node and the first two triangles of a leaf are both
// 96 bytes, so lanes at nodes and lanes at leaves (and, in the frame kernel,
// lanes completing a hit: a 48-byte record) share one fetch of six 16-byte
// words and one wait.  The triangle array carries one spare record, so a
// one-triangle leaf at its end reads in bounds.
__device__ inline const float4* step_fetch(const Trav& s) {
  return (s.cur & kLeafBit) ? s.tris + (size_t)(s.cur & kLeafIndexMask) * 3 : s.nodes + (size_t)s.cur * 6;
}

// One traversal step of an 8-wide tree on the six words step_fetch pointed to.
template <int D, bool COUNT, bool TOP = false>
__device__ inline bool trav_step_fetched(Trav& s, Hit& best, Stack<D>& st, Counters& c, const float4 (&v)[6]) {
  if (COUNT) c.steps++;
  if (s.cur & kLeafBit) {
    const uint32_t first = s.cur & kLeafIndexMask;
    const uint32_t count = ((s.cur >> kLeafCountShift) & 3u) + 1u;
    if (test_triangle<COUNT>(s, best, c, v[0], v[1], v[2])) return true;
    if (count > 1 && test_triangle<COUNT>(s, best, c, v[3], v[4], v[5])) return true;
    // leaves of more than two triangles (meshes built with leaf_size > 2)
    for (uint32_t k = 2; k < count; ++k) {
      const float4* tp = s.tris + (size_t)(first + k) * 3;
      if (test_triangle<COUNT>(s, best, c, gload4(tp), gload4(tp + 1), gload4(tp + 2))) return true;
    }
    s.cur = kNone;
  } else {
    s.cur = decode_node<8, D, COUNT>(v, s.r, s.thi, st, c, !s.any);
  }
  return s.cur == kNone ? pop_next<D, TOP>(s, st) : false;
}

template <int W, int D, bool COUNT, bool TOP = false>
__device__ inline bool trav_step(Trav& s, Hit& best, Stack<D>& st, Counters& c) {
  const bool leaf = (s.cur & kLeafBit) != 0u;
  const uint32_t first = s.cur & kLeafIndexMask;
  const uint32_t count = ((s.cur >> kLeafCountShift) & 3u) + 1u;
  if constexpr (W == 8) {
    const float4* p = step_fetch(s);
    float4 v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = gload4(p + k);
    return trav_step_fetched<D, COUNT, TOP>(s, best, st, c, v);
  } else {
    if (COUNT) c.steps++;
    if (leaf) {
      // all of the leaf's triangles are known up front: load them two at a time
      for (uint32_t k = 0; k < count; k += 2) {
        const float4* tp = s.tris + (size_t)(first + k) * 3;
        const bool two = k + 1 < count;
        float4 a0 = gload4(tp), b0 = gload4(tp + 1), c0 = gload4(tp + 2);
        float4 a1 = a0, b1 = b0, c1 = c0;
        if (two) a1 = gload4(tp + 3), b1 = gload4(tp + 4), c1 = gload4(tp + 5);
        if (test_triangle<COUNT>(s, best, c, a0, b0, c0)) return true;
        if (two && test_triangle<COUNT>(s, best, c, a1, b1, c1)) return true;
      }
      s.cur = kNone;
    } else {
      s.cur = visit_node<W, D, COUNT>(s.nodes, s.cur, s.r, s.thi, st, c);
    }
    return s.cur == kNone ? pop_next<D, TOP>(s, st) : false;
  }
}

// Wave helpers
__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ inline uint32_t rank_in_mask(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Frames split their tile sequence into kBands bands of consecutive
// kBandBlock-tile blocks, one per XCD: workgroups are dispatched to the eight
// XCDs round robin, so block b starts on band b % 8, and each XCD's L2
// serves the rays of one compact image band.  A wave whose band is drained
// moves on to the next band.  Each band has its own queue counter, 64 bytes
// apart, from word kBandQueue of the queue array.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kBandsPerXcd = 1;  // 2 and 4 (an XCD draining its bands in turn) measured slower
constexpr uint32_t kBands = kXcds * kBandsPerXcd;
constexpr uint32_t kBandBlock = 256;
constexpr uint32_t kBandQueue = 32;
constexpr uint32_t kQueueWords = kBandQueue + 16 * kBands;

// First ray id (tile position * 64) of band b of a frame of n_tiles tiles.
__host__ __device__ inline uint32_t band_first_ray(uint32_t b, uint32_t n_tiles) {
  const uint32_t nb = (n_tiles + kBandBlock - 1) / kBandBlock;
  const uint32_t t = (b * nb / kBands) * kBandBlock;
  return (t < n_tiles ? t : n_tiles) * 64u;
}

// Wave-local pool of ray ids taken from a global atomic queue in batches.
struct RayPool {
  uint32_t next, end;
  bool exhausted;
  uint32_t band, tried;  // banded frames: current band, bands drained so far
  __device__ void init() {
    next = end = 0;
    exhausted = false;
    band = (blockIdx.x % kXcds) * kBandsPerXcd;
    tried = 0;
  }
  // Refills [next, end) from the frame's bands (all lanes must call).
  __device__ void refill_banded(uint32_t* queue, uint32_t batch, uint32_t n_tiles) {
    while (tried < kBands) {
      const uint32_t lo = band_first_ray(band, n_tiles), hi = band_first_ray(band + 1, n_tiles);
      uint32_t base = 0;
      if (lane_id() == 0) base = atomicAdd(queue + kBandQueue + 16 * band, batch);
      base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
      if (base < hi - lo) {
        next = lo + base;
        end = min(next + batch, hi);
        return;
      }
      band = band + 1 == kBands ? 0 : band + 1;
      ++tried;
    }
    exhausted = true;
  }
  // As take, for a frame's banded tile sequence.
  __device__ uint32_t take_banded(bool want, uint32_t* queue, uint32_t batch, uint32_t n_tiles) {
    uint32_t id = kNone;
    uint64_t m = __ballot(want);
    while (m && !exhausted) {
      if (next >= end) {
        refill_banded(queue, batch, n_tiles);
        if (exhausted) break;
      }
      uint32_t avail = end - next;
      uint32_t r = rank_in_mask(m);
      bool mine = ((m >> lane_id()) & 1ull) && r < avail;
      if (mine) id = next + r;
      uint32_t n = (uint32_t)__popcll(m);
      uint32_t used = n < avail ? n : avail;
      next += used;
      m &= ~__ballot(mine);
    }
    return id;
  }
  // Hands ray ids to the lanes with want == true (all lanes must call).
  // Returns this lane's id, or kNone.
  __device__ uint32_t take(bool want, uint32_t* queue, uint32_t batch, uint32_t n_rays) {
    uint32_t id = kNone;
    uint64_t m = __ballot(want);
    while (m && !exhausted) {
      if (next >= end) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(queue, batch);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
        if (base >= n_rays) {
          exhausted = true;
          break;
        }
        next = base;
        end = min(base + batch, n_rays);
      }
      uint32_t avail = end - next;
      uint32_t r = rank_in_mask(m);
      bool mine = ((m >> lane_id()) & 1ull) && r < avail;
      if (mine) id = next + r;
      uint32_t n = (uint32_t)__popcll(m);
      uint32_t used = n < avail ? n : avail;
      next += used;
      m &= ~__ballot(mine);
    }
    return id;
  }
};

// ---- cooperative tail mode ---------------------------------------------------
// Once a wave's ray pool is exhausted and at most 8 rays remain in it, the
// remaining rays are each given a group of 8 lanes.  Every lane of a group
// holds an identical copy of the ray's state; per step lane j decodes child j
// of the node (or tests triangle j of the leaf) and the group combines the
// results with cross-lane shuffles.  This cuts the per-step instruction chain
// of the last, longest rays, which bound the kernel's duration.

__device__ inline uint32_t group_lane() { return threadIdx.x & 7u; }
__device__ inline uint32_t group_mask(uint64_t m) { return (uint32_t)(m >> (lane_id() & ~7u)) & 0xFFu; }

// Replaces v by lane `src`'s v, one 32-bit word at a time (no second copy live).
template <class T>
__device__ inline void shfl_from(T& v, int src) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words only");
  uint32_t* a = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) a[i] = (uint32_t)__shfl((int)a[i], src, 64);
}

// Cooperative node visit: lane j of the group handles child j.  Returns the
// nearest entered child (identical in all lanes of the group) and pushes the
// others far-to-near on the group's shared stack.
template <int W, int D, bool COUNT>
__device__ inline uint32_t coop_visit(const float4* __restrict__ nodes, uint32_t node, const Ray& r, float thi,
                                      Stack<D>& st, Counters& c) {
  constexpr int kVec = W == 8 ? 6 : 4;
  const float4* np = nodes + (size_t)node * kVec;
  const uint32_t j = group_lane();
  const float4 h = gload4(np);
  const float4 qx = gload4(np + 1), qy = gload4(np + 2), qz = gload4(np + 3);
  if (COUNT && j == 0) c.nodes++;
  const uint32_t ex = __float_as_uint(h.w);
  const float scx = __uint_as_float((ex & 0xFFu) << 23) * r.ix;
  const float scy = __uint_as_float(((ex >> 8) & 0xFFu) << 23) * r.iy;
  const float scz = __uint_as_float(((ex >> 16) & 0xFFu) << 23) * r.iz;
  const float bx = (h.x - r.ox) * r.ix, by = (h.y - r.oy) * r.iy, bz = (h.z - r.oz) * r.iz;
#define U(x) __float_as_uint(x)
  uint32_t lox, hix, loy, hiy, loz, hiz, ref;
  const int b = (int)(j & 3u);
  if (W == 4) {
    lox = U(qx.x), hix = U(qx.y), loy = U(qx.z), hiy = U(qx.w), loz = U(qy.x), hiz = U(qy.y);
    const float4 rr = gload4(np + 3);
    ref = j == 0 ? U(qy.z) : j == 1 ? U(qy.w) : j == 2 ? U(rr.x) : U(rr.y);
  } else {
    const bool up = j >= 4;
    lox = up ? U(qx.y) : U(qx.x), hix = up ? U(qx.w) : U(qx.z);
    loy = up ? U(qy.y) : U(qy.x), hiy = up ? U(qy.w) : U(qy.z);
    loz = up ? U(qz.y) : U(qz.x), hiz = up ? U(qz.w) : U(qz.z);
    const float4 rr = gload4(np + (up ? 5 : 4));
    ref = b == 0 ? U(rr.x) : b == 1 ? U(rr.y) : b == 2 ? U(rr.z) : U(rr.w);
  }
#undef U
  const float tlo = r.t_near * kSlackDown;
  const float lx = fmaf(byte_f(lox, b), scx, bx), ux = fmaf(byte_f(hix, b), scx, bx);
  const float ly = fmaf(byte_f(loy, b), scy, by), uy = fmaf(byte_f(hiy, b), scy, by);
  const float lz = fmaf(byte_f(loz, b), scz, bz), uz = fmaf(byte_f(hiz, b), scz, bz);
  const float tnx = r.ix < 0.f ? ux : lx, tfx = r.ix < 0.f ? lx : ux;
  const float tny = r.iy < 0.f ? uy : ly, tfy = r.iy < 0.f ? ly : uy;
  const float tnz = r.iz < 0.f ? uz : lz, tfz = r.iz < 0.f ? lz : uz;
  const float en = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tlo));
  const float exd = fminf(fminf(fminf(tfx, tfy), tfz) * kSlackUp, thi);
  float dist = (W == 8 || j < 4) && en <= exd ? en : INFINITY;
  // bitonic sort of the group's 8 (dist, ref) pairs, ascending
#pragma unroll
  for (int k = 2; k <= 8; k <<= 1) {
#pragma unroll
    for (int m = k >> 1; m >= 1; m >>= 1) {
      const float od = __shfl_xor(dist, m, 64);
      const uint32_t orf = (uint32_t)__shfl_xor((int)ref, m, 64);
      const bool asc = k == 8 || (j & (uint32_t)k) == 0;
      const bool low = (j & (uint32_t)m) == 0;
      const bool take = low == asc ? od < dist : od > dist;
      if (take) dist = od, ref = orf;
    }
  }
  const uint32_t hits = group_mask(__ballot(dist != INFINITY));
  const int n_hit = __popc(hits);
  const uint32_t first = (uint32_t)__shfl((int)ref, (int)(lane_id() & ~7u), 64);
  // lanes 1..n_hit-1 push, farthest deepest: lane j goes to entry sp + n_hit - 1 - j
  if (n_hit > 1) {
    if (j >= 1 && (int)j < n_hit) {
      st.put(st.sp + (uint32_t)(n_hit - 1) - j, ref, dist);
    }
    st.sp += (uint32_t)(n_hit - 1);
  }
  return n_hit ? first : kNone;
}

// Cooperative step: as trav_step, lanes of a group splitting the node's
// children or the leaf's triangles.  All lanes of the group end identical.
template <int W, int D, bool COUNT, bool TOP = false>
__device__ inline bool coop_step(Trav& s, Hit& best, Stack<D>& st, Counters& c) {
  const uint32_t j = group_lane();
  if (COUNT && j == 0) c.steps++;
  if (s.cur & kLeafBit) {
    const uint32_t first = s.cur & kLeafIndexMask;
    const uint32_t count = ((s.cur >> kLeafCountShift) & 3u) + 1u;
    float t = INFINITY, u = 0.f, v = 0.f;
    uint32_t id = kMiss;
    if (j < count) {
      const float4* tp = s.tris + (size_t)(first + j) * 3;
      const float4 a = gload4(tp), e1 = gload4(tp + 1), e2 = gload4(tp + 2);
      float tt, uu, vv;
      if (intersect_triangle(s.r.ox, s.r.oy, s.r.oz, s.r.dx, s.r.dy, s.r.dz, a.x, a.y, a.z, e1.x, e1.y, e1.z, e2.x,
                             e2.y, e2.z, s.r.t_near, s.tmax, tt, uu, vv))
        t = tt, u = uu, v = vv, id = __float_as_uint(a.w);
    }
    if (COUNT && j == 0) c.tris += count;
    // lexicographic (t, id) minimum over the group
#pragma unroll
    for (int m = 1; m <= 4; m <<= 1) {
      const float ot = __shfl_xor(t, m, 64), ou = __shfl_xor(u, m, 64), ov = __shfl_xor(v, m, 64);
      const uint32_t oid = (uint32_t)__shfl_xor((int)id, m, 64);
      if (ot < t || (ot == t && oid < id)) t = ot, u = ou, v = ov, id = oid;
    }
    if (id != kMiss) {
      if (s.any) {
        best.found = true;
        best.t = t, best.u = u, best.v = v, best.tri = id, best.obj = s.obj;
        return true;
      }
      bool better = !best.found || t < best.t ||
                    (t == best.t && (s.obj < best.obj || (s.obj == best.obj && id < best.tri)));
      if (better) {
        best.found = true;
        best.t = t, best.u = u, best.v = v, best.tri = id, best.obj = s.obj;
        s.tmax = t;
        s.thi = t * kSlackUp;
      }
    }
    s.cur = kNone;
  } else {
    s.cur = coop_visit<W, D, COUNT>(s.nodes, s.cur, s.r, s.thi, st, c);
  }
  return s.cur == kNone ? pop_next<D, TOP>(s, st) : false;
}

// ---- split tail mode ----------------------------------------------------------
// The alternative use of lane groups for the last rays of a wave: the lanes of
// a group (gsz = 2, 4 or 8 of them: 64 / gsz groups for the wave's remaining
// rays) each traverse their own part of the ray's tree with a stack of their
// own, so a long ray's chain of dependent node fetches runs up to gsz ways in
// parallel.  Group lane 0 keeps the ray's stack; the other lanes start idle,
// each with an empty stack in the LDS and spill column of a lane whose ray is
// complete.  After every step the idle lanes take the top entry of busy lanes'
// stacks, and the group's best t becomes every lane's culling bound.  A lane's
// hits stay its own until the group runs out of work; the group's result is
// then the least hit in the order (t, object, triangle), the order the
// single-lane traversal keeps, so it does not depend on which lane found what.
// (Scenes with a top level keep the cooperative tail.)

__device__ inline uint32_t split_lane(uint32_t gsz) { return lane_id() & (gsz - 1u); }
// bits of the lanes of this lane's group in a wave-wide mask
__device__ inline uint32_t split_mask(uint64_t m, uint32_t gsz) {
  return (uint32_t)((m >> (lane_id() & ~(gsz - 1u))) & ((1ull << gsz) - 1ull));
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ inline uint32_t nth_set(uint32_t m, uint32_t k) {
  for (uint32_t i = 0; i < k; ++i) m &= m - 1u;
  return (uint32_t)__builtin_ctz(m);
}

// best := the group's least hit in (t, object, triangle) order, in every lane
// (a hit found beats none).  Equal entries are the same hit: distinct hits of
// one object differ in their triangle.
__device__ inline void group_best(Hit& best, uint32_t gsz) {
  for (uint32_t m = 1; m < gsz; m <<= 1) {
    const float ot = __shfl_xor(best.t, (int)m, 64), ou = __shfl_xor(best.u, (int)m, 64);
    const float ov = __shfl_xor(best.v, (int)m, 64);
    const uint32_t otri = (uint32_t)__shfl_xor((int)best.tri, (int)m, 64);
    const uint32_t oobj = (uint32_t)__shfl_xor((int)best.obj, (int)m, 64);
    const bool ofound = __shfl_xor(best.found ? 1 : 0, (int)m, 64) != 0;
    const bool take = ofound && (!best.found || ot < best.t ||
                                 (ot == best.t && (oobj < best.obj || (oobj == best.obj && otri < best.tri))));
    if (take) {
      best.found = true;
      best.t = ot, best.u = ou, best.v = ov, best.tri = otri, best.obj = oobj;
    }
  }
}

// One step of every lane of a group in split mode (called by all gsz lanes).
// Returns true in every lane when the ray's traversal of the object is
// complete -- every lane idle with an empty stack, or, for an any-hit ray, a
// hit found by any lane -- and best then holds the group's result.
template <int W, int D, bool COUNT>
__device__ inline bool split_step(Trav& s, Hit& best, Stack<D>& st, Counters& c, uint32_t gsz) {
  bool busy = s.cur != kNone;
  if (busy && trav_step<W, D, COUNT>(s, best, st, c)) {
    busy = false;  // stack run dry (or an any-hit ray's hit)
    s.cur = kNone;
  }
  if (s.any && split_mask(__ballot(best.found), gsz)) {
    group_best(best, gsz);
    return true;
  }
  float bt = best.found ? best.t : INFINITY;
  for (uint32_t m = 1; m < gsz; m <<= 1) bt = fminf(bt, __shfl_xor(bt, (int)m, 64));
  if (bt < s.tmax) {
    s.tmax = bt;
    s.thi = bt * kSlackUp;
  }
  const uint32_t idle = split_mask(__ballot(!busy), gsz);
  const uint32_t donors = split_mask(__ballot(busy && st.sp > 0u), gsz);
  if (!idle) return false;
  if (!donors) {
    // all gsz lanes take part (a group's lanes change phase together); compared
    // with the lanes present, so that a group could never wait on an absent one
    if (idle != split_mask(__ballot(true), gsz)) return false;
    group_best(best, gsz);
    return true;
  }
  // Idle lanes are dealt round-robin over the donors: the k-th idle lane takes
  // entry k / n_d from the top of donor k mod n_d, reading it from the
  // donor's LDS column (a stack's top D entries are always there, and no donor
  // gives more than gsz - 1 < D).  Then each donor drops what it gave, moving
  // spilled entries back into the freed slots as pop does.
  const uint32_t j = split_lane(gsz), below = (1u << j) - 1u;
  const uint32_t n_d = (uint32_t)__popc(donors), n_i = (uint32_t)__popc(idle);
  int src = (int)lane_id();
  uint32_t k = 0, depth = 0;
  if (!busy) {
    k = (uint32_t)__popc(idle & below);
    depth = k / n_d;
    src = (int)((lane_id() & ~(gsz - 1u)) + nth_set(donors, k % n_d));
  }
  const uint32_t dsp = (uint32_t)__shfl((int)st.sp, src, 64);
  lds_entry* dl = st.lds;
  shfl_from(dl, src);
  j3d_u32x2 got = {kNone, 0u};
  const bool take = !busy && depth < dsp;
  if (take) got = dl[((dsp - 1u - depth) & (D - 1)) * st.stride];
  // every receiver has read its entry before any donor refills a slot
  __builtin_amdgcn_wave_barrier();
  if (busy && st.sp > 0u) {
    const uint32_t r = (uint32_t)__popc(donors & below);
    const uint32_t dealt = n_i > r ? (n_i - r + n_d - 1u) / n_d : 0u;
    for (uint32_t i = 0; i < min(dealt, st.sp); ++i) {
      --st.sp;
      if (st.sp >= (uint32_t)D) st.lds[(st.sp & (D - 1)) * st.stride] = st.spill[(uint64_t)(st.sp - D) * st.sstride];
    }
  }
  if (take && __uint_as_float(got.y) <= s.thi) s.cur = got.x;
  return false;
}

// ---- scenes of more than J3DG_MAX_OBJECTS objects: a top level ------------
//
// A binary BVH over the objects' world boxes (j3d's
// qbvh_two_level_with_transformations, canvas.cpp:762): node 0 is the root;
// a child ref with kLeafBit set is an object index.  Its entries share the
// lane's stack: an object's traversal starts with base = the stack depth, and
// pops stop there, so when it runs dry the top level's entries are next.  The
// top level is walked here, inline, when a ray starts and whenever an object
// is finished; lanes never fetch top-level nodes through the step machinery.
// World boxes are tested widened by `delta` (a bound on how far a hit found
// in an object's own frame can lie from the world ray, from rounding in the
// transforms) on top of the usual t slack, so no object with a hit in range
// is culled; results are independent of the order objects are visited in.
constexpr uint32_t kTopGuard = 1u << 22;

struct TopNode {
  float box[2][6];  // the children's world boxes: lo xyz, hi xyz
  uint32_t child[2];  // kNone for an empty slot (a one-object scene)
  uint32_t pad[2];
};

// The top level is read from global memory, or from a block's LDS copy when
// it is small enough (both through typed pointers: a generic pointer would
// compile to flat loads).
typedef const __attribute__((address_space(3))) TopNode lds_top;
__device__ inline float top_box(const TopNode* n, int k, int a) { return gload(&n->box[k][a]); }
__device__ inline float top_box(lds_top* n, int k, int a) { return n->box[k][a]; }
__device__ inline uint32_t top_child(const TopNode* n, int k) { return gload(&n->child[k]); }
__device__ inline uint32_t top_child(lds_top* n, int k) { return n->child[k]; }

// Entry distance of the world ray into child k's box widened by delta, or
// INFINITY when it misses [tlo, thi].
template <class NP>
__device__ inline float top_entry(NP n, int k, const Ray& w, float delta, float tlo, float thi) {
  const float o[3] = {w.ox, w.oy, w.oz}, inv[3] = {w.ix, w.iy, w.iz};
  float en = tlo, ex = thi;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float p = (top_box(n, k, a) - delta - o[a]) * inv[a];
    const float q = (top_box(n, k, 3 + a) + delta - o[a]) * inv[a];
    en = fmaxf(en, fminf(p, q));
    ex = fminf(ex, fmaxf(p, q) * kSlackUp);
  }
  return en <= ex ? en : INFINITY;
}

__device__ inline Ray world_ray(f4 o, f4 d, float tn, float tf) {
  Ray r;
  r.ox = o.x, r.oy = o.y, r.oz = o.z;
  r.dx = d.x, r.dy = d.y, r.dz = d.z;
  r.ix = safe_rcp(d.x), r.iy = safe_rcp(d.y), r.iz = safe_rcp(d.z);
  r.t_near = tn;
  r.t_far = tf;
  return r;
}

// Walks the top level from s.cur (a top node, or kNone: pop the top level's
// entries) to the next object whose box the ray enters within the current
// bound, and starts that object's traversal (s.cur = its root).  Returns true
// when no object is left: the ray is done.  The lane's stack holds top-level
// entries only when this is called.
template <int D, class NP>
__device__ inline bool top_next(Trav& s, Stack<D>& st, const DevObject* __restrict__ objs, NP top, float delta,
                                f4 wo, f4 wd) {
  if (!ray_finite(wo, wd)) {
    s.cur = kNone;
    return true;
  }
  const Ray w = world_ray(wo, wd, s.r.t_near, s.r.t_far);
  const float tlo = w.t_near * kSlackDown;
  uint32_t cur = s.cur;
  // each trip pops, enters an object or descends one level of a proper tree;
  // the bound only guards against a malformed one (reported as an overflow)
  for (uint32_t guard = 0;; ++guard) {
    if (guard == kTopGuard) {
      st.overflow = true;
      s.cur = kNone;
      return true;
    }
    if (cur == kNone) {
      uint32_t ref;
      float dist;
      bool found = false;
      while (st.pop(ref, dist)) {
        if (dist <= s.thi) {
          cur = ref;
          found = true;
          break;
        }
      }
      if (!found) {
        s.cur = kNone;
        return true;
      }
    }
    if (cur & kLeafBit) {
      const uint32_t oi = cur & kLeafIndexMask;
      const DevObject& ob = objs[oi];
      s.r = make_object_ray(ob.inv_cs, wo, wd, w.t_near, w.t_far);
      s.nodes = ob.nodes;
      s.tris = ob.tris;
      s.obj = oi;
      s.base = st.sp;
      s.cur = 0;  // the object's root, an internal node
      return false;
    }
    const NP n = top + cur;
    const uint32_t c0 = top_child(n, 0), c1 = top_child(n, 1);
    float e0 = c0 != kNone ? top_entry(n, 0, w, delta, tlo, s.thi) : INFINITY;
    float e1 = c1 != kNone ? top_entry(n, 1, w, delta, tlo, s.thi) : INFINITY;
    uint32_t r0 = c0, r1 = c1;
    if (e1 < e0) {
      const float te = e0;
      e0 = e1, e1 = te;
      r0 = c1, r1 = c0;
    }
    if (e1 != INFINITY) st.push(r1, e1);
    cur = e0 != INFINITY ? r0 : kNone;
  }
}

}  // namespace j3d
