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
#include <functional>
#include <chrono>
#include <cmath>
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
#include "traverse.h"

using namespace j3d;

// ============================================================ device side ===

namespace j3d {

// Event bits a lane raises during one loop trip; the wave turns them into
// uniform counts with ballot + popcount (no per-lane counters).
enum : uint32_t {
  kEvPrimary = 1u, kEvHit = 2u, kEvShadowed = 4u, kEvResolved = 8u, kEvFallback = 16u, kEvShadowRay = 32u
};

// Traversal steps a lane may take in one wave loop trip, measured on the
// 28M-triangle frame: with the cooperative tail 3 was best (1 -> 0.87 ms,
// 2 -> 0.84, 3 -> 0.80, 4 -> 0.82); with the split tail 4 is (2 -> 0.723-0.728,
// 3 -> 0.697-0.707, 4 -> 0.679-0.697, 5 -> 0.708-0.712, 6 -> 0.705-0.707).
constexpr int kStepsPerTrip = 4;
// A cooperative group (the tail with at most 8 rays left in the wave) takes up
// to kCoopSteps steps per trip: no refill shares its trips any more (28M-
// triangle frame: 1 -> 0.80 ms, 4 -> 0.78; 8 traversal steps per trip for the
// tail's lanes outside groups, 0.79, and 16, 0.81, did not pay).
constexpr int kCoopSteps = 4;

// Lane phases of k_trace.
enum : uint32_t { kIdle = 0u, kBounded = 1u, kPlain = 2u, kShadowPhase = 3u, kFinishPhase = 4u };

// The first G-buffer word of a hit (j3d/canvas.cpp:801-834): mark 0, or mark
// 2 with r, g, b in bytes 1-3 when the object has a surface colour under the
// frame's flags (texture first, then vertex colours).  Sets `colored`.
__device__ inline uint32_t hit_color_word(const DevObject& ob, uint32_t flags, uint32_t tri, float u, float v,
                                          bool& colored) {
  const float s0 = 1.f - u - v;
  colored = false;
  if ((flags & J3DG_TEXTURED) && ob.uv) {
    const float* q = ob.uv + 6 * (size_t)tri;
    float cx = s0 * gload(q) + u * gload(q + 2) + v * gload(q + 4);
    float cy = s0 * gload(q + 1) + u * gload(q + 3) + v * gload(q + 5);
    // std::max(std::min(c, 1), 0) exactly, NaN included
    cx = (1.f < cx) ? 1.f : cx;
    cx = (cx < 0.f) ? 0.f : cx;
    cy = (1.f < cy) ? 1.f : cy;
    cy = (cy < 0.f) ? 0.f : cy;
    const int w = (int)ob.tex_w, h = (int)ob.tex_h;
    const float fx = cx * (float)w, fy = cy * (float)h;
    int x = fx == fx ? (int)fx : -1;  // NaN: below 0, clamped like an x86 conversion
    int y = fy == fy ? (int)fy : -1;
    x = x < 0 ? 0 : x >= w ? w - 1 : x;
    y = y < 0 ? 0 : y >= h ? h - 1 : y;
    const uint32_t color = gload(ob.tex + (size_t)y * ob.tex_w + x);
    colored = true;
    return 2u | (color & 0xFFu) << 8 | ((color >> 8) & 0xFFu) << 16 | ((color >> 16) & 0xFFu) << 24;
  }
  if ((flags & J3DG_VERTEXCOLORS) && ob.vcol) {
    const uint32_t* tr = ob.idx + 3 * (size_t)tri;
    const float* c0 = ob.vcol + 3 * (size_t)gload(tr);
    const float* c1 = ob.vcol + 3 * (size_t)gload(tr + 1);
    const float* c2 = ob.vcol + 3 * (size_t)gload(tr + 2);
    uint32_t word = 2u;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      word |= color_byte(gload(c0 + k) * s0 + u * gload(c1 + k) + v * gload(c2 + k)) << (8 * (k + 1));
    colored = true;
    return word;
  }
  return 0u;
}

// Writes the first G-buffer word of a hit: the mark byte alone (r, g, b
// untouched, j3d/canvas.cpp:801) or mark and colour together.
__device__ inline void store_hit_word(char* pp, uint32_t word, bool colored) {
  if (colored)
    *reinterpret_cast<uint32_t*>(pp) = word;
  else
    *reinterpret_cast<uint8_t*>(pp) = 0;
}

// A primary hit's remaining work (j3d/canvas.cpp:788-801): view-space normal
// object_cs * (CS^-1 * n_tri), mark 0, db_id; and, for shadows, the world-space
// hit point from the object's vertices and the direction to the light
// (canvas.cpp:836-853).  r0..r2 is the triangle's completion record
// (DevObject::fin): its vertices and normal.  A frame with surface colours
// (canvas.cpp:803-834) gets them from k_color after the trace: here the
// db_id word then carries the object index for it (the colour code stays out
// of the traversal loop, where it cost registers and time even untaken).
// Returns whether a shadow ray follows.
template <bool COLOR>
__device__ inline bool finish_hit(const FrameParams& f, const DevObject& ob, const Hit& best, bool writer, char* pp,
                                  const float4& r0, const float4& r1, const float4& r2, f4& pos, f4& ld) {
  f4 n = {r0.w, r1.w, r2.w, 0.f};
  n = mat_vec(f.cs_inv, n);
  n = mat_vec(ob.cs, n);
  float* pf = reinterpret_cast<float*>(pp);
  uint32_t* pu = reinterpret_cast<uint32_t*>(pp);
  if (writer) {
    *reinterpret_cast<uint8_t*>(pp) = 0;  // mark; r, g, b untouched (j3d/canvas.cpp:801)
    pf[1] = n.x;
    pf[2] = n.y;
    pf[3] = best.t;
    pu[4] = best.tri;
    pf[5] = best.u;
    pf[6] = best.v;
    pu[7] = COLOR ? best.obj : ob.db_id;
  }
  if (!(f.flags & J3DG_SHADOW)) return false;
  f4 V[3];
  V[0] = mat_vec(ob.cs, f4{r0.x, r0.y, r0.z, 1.f});
  V[1] = mat_vec(ob.cs, f4{r1.x, r1.y, r1.z, 1.f});
  V[2] = mat_vec(ob.cs, f4{r2.x, r2.y, r2.z, 1.f});
  const float bu = best.u, bv = best.v, s0 = 1.f - bu - bv;
  pos = f4{V[0].x * s0 + bu * V[1].x + bv * V[2].x, V[0].y * s0 + bu * V[1].y + bv * V[2].y,
           V[0].z * s0 + bu * V[1].z + bv * V[2].z, V[0].w * s0 + bu * V[1].w + bv * V[2].w};
  if (f.flags & J3DG_DIRECTIONAL_LIGHT)
    ld = f4{f.light_dir[0], f.light_dir[1], f.light_dir[2], 0.f};
  else
    ld = f4{f.light[0] - pos.x, f.light[1] - pos.y, f.light[2] - pos.z, f.light[3] - pos.w};
  return true;
}

__device__ inline void add_counter(uint64_t* counters, int k, uint64_t v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane_id() == 0 && v) atomicAdd((unsigned long long*)&counters[k], (unsigned long long)v);
}

// Persistent waves with per-lane refill.  Counters (uint64):
//   0 primary rays  1 hits  2 shadow rays  3 shadowed  4 node visits  5 triangle tests
//   6 stack overflows  7 learned: resolved by the bounded pass  8 learned: fell back
//   9 wave loop trips  10 lane steps
//
// kModeFrame: primary rays of FrameParams' rectangle; a miss writes the miss
//   fields (j3d/canvas.cpp:859-866).  With f.fused a hit is completed in the
//   lane (finish_hit) and its shadow ray traced right after by the same lane,
//   so the frame is one launch; otherwise a hit writes depth, object_id and
//   the barycentrics and parks the object index in db_id for k_finish.
// kModeShadow: any-hit rays from the shadow queue k_finish filled; a hit sets
//   the pixel's mark bit 0 (j3d/canvas.cpp:855-856).
// kModeRays / kModeAnyRays: world-space rays from a buffer (ray queries).
// kModeBatch: one-launch frames of a batch; a lane looks up its tile's frame
//   (binary search over the frames' first tiles) and reads that frame's
//   parameters from the device table.
template <int MODE>
__device__ inline const FrameParams& frame_ref(const FrameParams& f, const BatchParams& bp, uint32_t fi) {
  if constexpr (MODE == kModeBatch)
    return bp.frames[fi];
  else
    return f;
}

// At least 3 waves per SIMD (at most 168 VGPRs): the tail modes' code would
// otherwise take the kernel just past that line, to 2 waves.  The counting
// instances (diagnostics) keep whatever registers they need.
template <int W, int D, bool COUNT, int MODE>
__global__ void __launch_bounds__(256, COUNT ? 1 : 3) k_trace(FrameParams f, RayQuery q, uint32_t n_rays_q,
                                               const DevObject* __restrict__ objs, uint32_t n_obj, uint2* spill,
                                               uint32_t spill_cap, uint32_t* queue, uint64_t* counters,
                                               uint32_t batch, uint32_t refill_min, uint32_t coop_tail,
                                               uint64_t* timeline, BatchParams bp) {
  constexpr int M = MODE & (kModeTop - 1);
  constexpr bool kTop = (MODE & kModeTop) != 0;  // more objects than the LDS table: a top-level BVH
  constexpr bool kIsFrame = M == kModeFrame || M == kModeFrameColor;
  constexpr bool kFrameLike = kIsFrame || M == kModeBatch;
  extern __shared__ uint2 lds_stack[];
  // the object table (at most J3DG_MAX_OBJECTS entries outside batches) is read
  // on every ray start and hit completion: keep the block's copy in LDS
  constexpr bool kLdsObjects = M != kModeBatch && !kTop;
  __shared__ DevObject s_objs[kLdsObjects ? J3DG_MAX_OBJECTS : 1];
  if (kLdsObjects) {
    n_obj = min(n_obj, (uint32_t)J3DG_MAX_OBJECTS);
    const uint32_t words = n_obj * (uint32_t)(sizeof(DevObject) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(objs);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s_objs);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = gload(src + i);
    __syncthreads();
  }
  const DevObject* const tab = kLdsObjects ? s_objs : objs;
  const TopNode* const top = kTop ? static_cast<const TopNode*>(f.top_nodes) : nullptr;
  // a small top level is copied into LDS after the stack (all threads reach the barrier)
  lds_top* top_l = nullptr;
  if (kTop && f.top_lds) {
    const uint32_t words = f.top_count * (uint32_t)(sizeof(TopNode) / 4);
    const uint32_t* src = static_cast<const uint32_t*>(f.top_nodes);
    uint32_t* dst = reinterpret_cast<uint32_t*>(lds_stack + D * blockDim.x);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = gload(src + i);
    __syncthreads();
    top_l = (lds_top*)(lds_stack + D * blockDim.x);
  }
  Stack<D> st;
  st.init(lds_stack, spill, spill_cap);
  Counters c = {0, 0, 0}, cs = {0, 0, 0};  // primary (or query) / shadow-phase work
  uint32_t acc[6] = {0, 0, 0, 0, 0, 0};  // wave-uniform event counts
  uint32_t trips = 0;
  uint32_t ray_steps = 0;            // counting: steps of the current ray
  uint32_t max_steps = 0, hist[7] = {0, 0, 0, 0, 0, 0, 0};
  uint32_t max_steps_s = 0, hist_s[7] = {0, 0, 0, 0, 0, 0, 0};  // fused shadow phase
  bool coop = false;                 // cooperative tail mode (wave-uniform)
  bool split = false;                // its split form: a group's lanes traverse parts of the ray's tree
  uint32_t gsz = 8;                  // lanes per group (wave-uniform; 2, 4 or 8 in split mode)
  // counting: per-wave timeline (shader clock at start, pool exhausted, tail mode, end; trips)
  uint64_t tl_start = COUNT ? (uint64_t)wall_clock64() : 0, tl_exh = 0, tl_coop = 0;
  uint32_t trips_coop = 0;
  RayPool pool;
  pool.init();
  uint32_t n_rays = n_rays_q;
  if (kIsFrame) n_rays = f.n_tiles * 64u;
  if (M == kModeShadow) n_rays = __builtin_amdgcn_readfirstlane(*q.count);
  const bool learned = kIsFrame && (f.flags & J3DG_LEARNED) && f.proposal;
  const bool any_mode = M == kModeAnyRays || M == kModeShadow;
  const bool fused = M == kModeBatch || (kIsFrame && f.fused);

  // lane state
  uint32_t phase = 0;  // 0 idle, 1 bounded (learned), 2 plain, 3 shadow (one-launch frames)
  uint32_t pix = 0;    // pixel index (frames, shadow queue) or ray id (queries)
  uint32_t cost_slot = 0, pix_trips = 0;  // frames: where this pixel's loop trips go, and their count
  uint32_t fi = 0;                        // batches: the lane's frame
  uint32_t lnobj = n_obj;                 // objects of the lane's frame
  const DevObject* lobjs = tab;           // and their table
  f4 wo = {0.f, 0.f, 0.f, 1.f}, wd = {0.f, 0.f, 0.f, 0.f};
  Trav s;
  s.any = any_mode;
  s.cur = kNone;
  s.obj = 0;
  Hit best;
  best.found = false;
  best.t = best.u = best.v = 0.f;
  best.tri = best.obj = kMiss;
  // starts the lane's ray (world o, d in wo, wd) over [tn, tf] at its first object
  auto start_ray = [&](float tn, float tf) {
    if constexpr (kTop) {
      s.r.t_near = tn;
      s.r.t_far = tf;
      s.tmax = best.found ? fminf(best.t, tf) : tf;
      s.thi = s.tmax * kSlackUp;
      s.base = 0;
      s.cur = 0;  // the top level's root; kNone after this if no object is entered
      if (top_l)
        top_next<D>(s, st, lobjs, top_l, f.top_delta, wo, wd);
      else
        top_next<D>(s, st, lobjs, top, f.top_delta, wo, wd);
    } else {
      begin_object(s, lobjs, 0, wo, wd, tn, tf, best);
    }
  };
  // an object's traversal reported done: unless an any-hit ray has its hit,
  // the top level supplies the next object
  auto resume_top = [&](bool done) {
    if (!kTop || !done || (s.any && best.found)) return done;
    s.cur = kNone;
    return top_l ? top_next<D>(s, st, lobjs, top_l, f.top_delta, wo, wd)
                 : top_next<D>(s, st, lobjs, top, f.top_delta, wo, wd);
  };
  // the tail modes: this lane takes over lane src's ray (all of its state)
  auto take_state = [&](int src) {
    shfl_from(phase, src);
    shfl_from(pix, src);
    shfl_from(ray_steps, src);
    shfl_from(cost_slot, src);
    shfl_from(pix_trips, src);
    if (M == kModeBatch) {  // the ray's frame travels with it
      shfl_from(fi, src);
      shfl_from(lnobj, src);
      shfl_from(lobjs, src);
    }
    shfl_from(wo, src);
    shfl_from(wd, src);
    shfl_from(s, src);
    shfl_from(best, src);
    shfl_from(st, src);
  };
  // split mode: a lane that starts idle, with an empty stack in the column of
  // the k-th lane of `free_cols` (lanes of this wave whose columns nobody owns)
  auto fresh_helper = [&](uint64_t free_cols, uint32_t k) {
    for (uint32_t i = 0; i < k; ++i) free_cols &= free_cols - 1u;
    st.init(lds_stack, spill, spill_cap, (threadIdx.x & ~63u) + (uint32_t)__builtin_ctzll(free_cols));
    s.cur = kNone;
  };

  for (;;) {
    uint32_t ev = 0;
    uint32_t n_idle = (uint32_t)__popcll(__ballot(phase == 0));
    if (n_idle == 64 && pool.exhausted) break;
    if (!pool.exhausted && (n_idle >= refill_min || n_idle == 64)) {
      // hand new rays to idle lanes; ids outside the rectangle leave a lane idle, so repeat
      while (__ballot(phase == 0) && !pool.exhausted) {
        uint32_t id = kIsFrame ? pool.take_banded(phase == 0, queue, batch, f.n_tiles)
                                         : pool.take(phase == 0, queue, batch, n_rays);
        if (id == kNone) continue;
        float tn, tf;
        if (kFrameLike) {
          uint32_t px, py;
          uint32_t tile = id >> 6;
          if (M == kModeBatch) {
            uint32_t lo = 0, hi = bp.n_frames;  // first_tile[lo] <= tile < first_tile[hi]
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (gload(bp.first_tile + mid) <= tile) lo = mid; else hi = mid;
            }
            fi = lo;
            tile -= gload(bp.first_tile + fi);
          }
          const FrameParams& F = frame_ref<M>(f, bp, fi);
          if (M == kModeBatch) {
            lnobj = F.n_obj;
            lobjs = objs + F.obj_base;
          }
          if (kIsFrame && f.tile_order) tile = gload(f.tile_order + tile);
          cost_slot = tile * 64u + (id & 63u);
          pix_trips = 0;
          if (!tile_pixel(F, tile, id & 63u, px, py)) {
            if (kIsFrame && f.pixel_cost) f.pixel_cost[cost_slot] = 0;
            continue;
          }
          primary_ray(F, px, py, wo, wd);
          tn = F.diagonal / 100.f;
          tf = FLT_MAX;
          pix = py * F.w + px;
          phase = 2;
          // a batch's learned frames carry their own proposal buffers
          const float2* prop = M == kModeBatch ? F.proposal : learned ? f.proposal : nullptr;
          if (prop) {
            const float2 pr = prop[pix];
            if (pr.y > 0.f) {  // predicted hit: bounded pass first (pr.x = t_pred * (1 + margin))
              tf = pr.x;
              phase = 1;
            }
          }
          ev |= kEvPrimary;
        } else if (M == kModeShadow) {
          const float4* rr = reinterpret_cast<const float4*>(q.rays) + 3 * (size_t)id;
          float4 o4 = gload4(rr), d4 = gload4(rr + 1), p4 = gload4(rr + 2);
          wo = f4{o4.x, o4.y, o4.z, o4.w};
          wd = f4{d4.x, d4.y, d4.z, d4.w};
          tn = 1e-3f;
          tf = FLT_MAX;
          pix = __float_as_uint(p4.x);
          phase = 2;
        } else {
          const float* rr = q.rays + 8 * (size_t)id;
          wo = f4{rr[0], rr[1], rr[2], 1.f};
          wd = f4{rr[3], rr[4], rr[5], 0.f};
          tn = rr[6];
          tf = rr[7];
          pix = id;
          phase = 2;
        }
        best.found = false;
        st.sp = 0;
        s.any = any_mode;
        if (COUNT) ray_steps = 0;
        if (lnobj) {
          start_ray(tn, tf);
        } else {
          s.cur = kNone;
          s.r.t_near = tn;
          s.r.t_far = tf;
        }
      }
    }
    if (kFrameLike && __ballot(ev & kEvPrimary)) acc[0] += (uint32_t)__popcll(__ballot(ev & kEvPrimary));
    ev &= ~kEvPrimary;
    if (COUNT && pool.exhausted && !tl_exh) tl_exh = (uint64_t)wall_clock64();
    if (coop_tail && !coop && pool.exhausted) {
      const uint64_t act = __ballot(phase != 0);
      const uint32_t n_act = (uint32_t)__popcll(act);
      // the split tail (coop_tail 2, 3, 4) starts with 8, 16 or 32 rays left, in
      // groups of 64 / (rays left rounded up to 8, 16, 32) lanes; the
      // cooperative tail and scenes with a top level wait for 8 rays
      const uint32_t most = (!kTop && coop_tail >= 2u) ? 8u << (coop_tail - 2u) : 8u;
      if (n_act > 0 && n_act <= most) {
        // hand each remaining ray to a group of gsz lanes (group g: the g-th active lane)
        gsz = n_act <= 8u ? 8u : n_act <= 16u ? 4u : 2u;
        const uint32_t g = lane_id() / gsz;
        uint64_t m = act;
        for (uint32_t i = 0; i < g && m; ++i) m &= m - 1;
        const bool has = g < n_act;
        take_state(has ? (int)__builtin_ctzll(m) : (int)lane_id());
        if (!has) phase = 0u, pix_trips = 0u;
        coop = true;
        split = !kTop && coop_tail >= 2u;
        // lanes 1 .. gsz-1 of group g take the free columns (gsz-1)g ..
        // (gsz-1)g + gsz-2 in lane order (the n_act <= 64 / gsz active lanes
        // leave at least (gsz - 1) n_act)
        if (split && has && split_lane(gsz) != 0u) fresh_helper(~act, g * (gsz - 1u) + split_lane(gsz) - 1u);
        if (COUNT) tl_coop = (uint64_t)wall_clock64();
      }
    }
    // split mode: as rays finish, the remaining ones move into wider groups
    // (64 / (rays left rounded up to 8, 16, 32) lanes).  A ray's lanes keep
    // their traversals and stacks (a lane's Stack points at the column it
    // owns); the new lanes start idle in columns that no remaining ray's lane
    // owns.  Groups stay at 8 lanes below 8 rays: 16 or 32 lanes per ray ran
    // slower (more speculative fetches per ray than its tree can use).
    if (split && gsz < 8u) {
      const uint64_t heads = __ballot(phase != 0u && split_lane(gsz) == 0u);
      const uint32_t n = (uint32_t)__popcll(heads);
      const uint32_t ng = n <= 8u ? 8u : n <= 16u ? 4u : 2u;
      if (n > 0u && ng > gsz) {
        const uint32_t col = (uint32_t)(st.lds - (lds_entry*)lds_stack) & 63u;
        uint64_t used = 0;
        for (uint32_t b = 0; b < 64u; ++b)
          if (__ballot(phase != 0u && col == b)) used |= 1ull << b;
        const uint32_t r = lane_id() / ng, jn = split_lane(ng);
        uint64_t m = heads;
        for (uint32_t i = 0; i < r && m; ++i) m &= m - 1u;
        const bool has = r < n;
        const uint32_t head = has ? (uint32_t)__builtin_ctzll(m) : lane_id();
        take_state((int)(has ? head + (jn < gsz ? jn : 0u) : lane_id()));
        if (!has) phase = 0u, pix_trips = 0u;
        // new lanes gsz .. ng-1 of group r: free columns r (ng - gsz) + ... in
        // lane order (n <= 64 / ng remaining rays own n gsz columns)
        if (has && jn >= gsz) fresh_helper(~used, r * (ng - gsz) + jn - gsz);
        gsz = ng;
      }
    }
    const bool writer = !coop || (lane_id() & (gsz - 1u)) == 0u;
    if (COUNT) trips++;
    if (COUNT && coop) trips_coop++;
    if (phase != 0) {
      if (COUNT) ray_steps++;
      if (kIsFrame) pix_trips++;
      Counters& cc = (kFrameLike && phase == kShadowPhase) ? cs : c;
      bool done;
      if (lnobj == 0 || (kTop && phase != kFinishPhase && s.cur == kNone)) {
        done = true;  // no object, or (top level) no object box on the ray's way or a ray that is not finite
      } else if (kFrameLike && (phase == kFinishPhase || (W == 8 && !coop))) {
        // one fetch for every lane: a node, a leaf's triangles or a hit's completion record
        const float4* p = phase == kFinishPhase ? lobjs[best.obj].fin + 3 * (size_t)best.tri : step_fetch(s);
        float4 v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = gload4(p + k);
        if (phase == kFinishPhase) {
          done = false;
          phase = 0;
          const FrameParams& F = frame_ref<M>(f, bp, fi);
          char* pp = reinterpret_cast<char*>(F.gbuf) + (size_t)pix * 32;
          f4 pos, ld;
          if (finish_hit<M == kModeFrameColor>(F, lobjs[best.obj], best, writer, pp, v[0], v[1], v[2], pos, ld)) {
            // the same lane traces the pixel's shadow ray next (any hit, t in [1e-3, inf))
            ev |= kEvShadowRay;
            phase = kShadowPhase;
            wo = pos;
            wd = ld;
            best.found = false;
            st.sp = 0;
            s.any = true;
            if (COUNT) ray_steps = 0;
            start_ray(1e-3f, FLT_MAX);
            if (split && split_lane(gsz) != 0u) s.cur = kNone;  // split mode: lane 0 starts, the others take work
          }
        } else {
          if constexpr (W == 8) {
            // up to kStepsPerTrip traversal steps per loop trip: the trip's
            // fixed work (refills, hit completions, bookkeeping: most of its
            // instructions) is shared by more steps
            done = resume_top(trav_step_fetched<D, COUNT, kTop>(s, best, st, cc, v));
#pragma unroll
            for (int extra = 1; extra < kStepsPerTrip; ++extra) {
              if (done) break;
              const float4* p2 = step_fetch(s);
#pragma unroll
              for (int k = 0; k < 6; ++k) v[k] = gload4(p2 + k);
              done = resume_top(trav_step_fetched<D, COUNT, kTop>(s, best, st, cc, v));
            }
          }
        }
      } else if (split) {
        // a group's lanes share `done`: they leave the loop together
        done = split_step<W, D, COUNT>(s, best, st, cc, gsz);
        for (int extra = 1; extra < kCoopSteps && !done; ++extra) done = split_step<W, D, COUNT>(s, best, st, cc, gsz);
      } else if (coop) {
        // a group's lanes share `done`: they leave the loop together
        done = resume_top(coop_step<W, D, COUNT, kTop>(s, best, st, cc));
        for (int extra = 1; extra < kCoopSteps && !done; ++extra)
          done = resume_top(coop_step<W, D, COUNT, kTop>(s, best, st, cc));
      } else {
        done = resume_top(trav_step<W, D, COUNT, kTop>(s, best, st, cc));
      }
      if (!kTop && done && !(s.any && best.found) && s.obj + 1 < lnobj) {
        begin_object(s, lobjs, s.obj + 1, wo, wd, s.r.t_near, s.r.t_far, best);
        st.sp = 0;
        done = false;
        if (split && split_lane(gsz) != 0u) s.cur = kNone;
      }
      if (done && COUNT && phase != kBounded && writer) {
        // per-ray step histograms (primary / shadow), kept in registers and reduced once at exit
        const uint32_t b = 31u - __clz(ray_steps | 1u);
        if (phase == kShadowPhase) {
          max_steps_s = max(max_steps_s, ray_steps);
          hist_s[b > 6 ? 6 : b]++;
        } else {
          max_steps = max(max_steps, ray_steps);
          hist[b > 6 ? 6 : b]++;
        }
      }
      if (done) {
        if (M == kModeAnyRays) {
          if (writer) q.any[pix] = best.found ? 1 : 0;
          phase = 0;
        } else if (M == kModeRays) {
          if (writer) {
            q.obj[pix] = best.found ? best.obj : kMiss;
            q.tri[pix] = best.found ? best.tri : kMiss;
            q.t[pix] = best.found ? best.t : FLT_MAX;
            q.u[pix] = best.found ? best.u : 0.f;
            q.v[pix] = best.found ? best.v : 0.f;
          }
          phase = 0;
        } else if (M == kModeShadow || phase == kShadowPhase) {
          if (best.found) {
            // mark |= 1 (j3d/canvas.cpp:855-856): bit 0 of the pixel's first word, colour bits kept
            if (writer)
              atomicOr(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(frame_ref<M>(f, bp, fi).gbuf) +
                                                   (size_t)pix * 32),
                       1u);
            ev |= kEvShadowed;
          }
          phase = 0;
        } else if (phase == 1 && !best.found) {
          // The bounded pass tested every triangle with t in [t_near, bound] and
          // found none, so the closest hit lies beyond the bound: the fallback
          // searches (bound, inf) only.  A bound that is NaN or below t_near
          // proves nothing and the fallback starts at t_near.
          ev |= kEvFallback;
          phase = 2;
          st.sp = 0;
          const float t0 = frame_ref<M>(f, bp, fi).diagonal / 100.f, tb = s.r.t_far;
          start_ray(tb >= t0 ? tb : t0, FLT_MAX);
          if (split && split_lane(gsz) != 0u) s.cur = kNone;
        } else {
          if (phase == 1) ev |= kEvResolved;
          phase = 0;
          const FrameParams& F = frame_ref<M>(f, bp, fi);
          char* pp = reinterpret_cast<char*>(F.gbuf) + (size_t)pix * 32;
          float* pf = reinterpret_cast<float*>(pp);
          uint32_t* pu = reinterpret_cast<uint32_t*>(pp);
          if (best.found && fused) {
            // completed on the lane's next trip, whose fetch brings the triangle's record
            ev |= kEvHit;
            phase = kFinishPhase;
          } else if (best.found) {
            ev |= kEvHit;
            if (writer) {
              pf[3] = best.t;
              pu[4] = best.tri;
              pf[5] = best.u;
              pf[6] = best.v;
              pu[7] = best.obj;  // object index; k_finish replaces it with the db_id
            }
          } else if (writer) {
            pf[1] = 0.f;
            pf[2] = 0.f;
            pf[3] = FLT_MAX;
            pu[4] = kMiss;
            pu[7] = 0u;
          }
        }
      }
    }
    if (kIsFrame && f.pixel_cost && phase == 0 && pix_trips && writer) {
      f.pixel_cost[cost_slot] = (uint16_t)min(pix_trips, 65535u);  // the pixel is complete
      pix_trips = 0;
    }
    if (!writer) ev = 0;
    if ((kFrameLike || M == kModeShadow) && __ballot(ev != 0)) {
      acc[0] += (uint32_t)__popcll(__ballot(ev & kEvPrimary));
      acc[1] += (uint32_t)__popcll(__ballot(ev & kEvHit));
      acc[2] += (uint32_t)__popcll(__ballot(ev & kEvShadowed));
      acc[3] += (uint32_t)__popcll(__ballot(ev & kEvResolved));
      acc[4] += (uint32_t)__popcll(__ballot(ev & kEvFallback));
      acc[5] += (uint32_t)__popcll(__ballot(ev & kEvShadowRay));
    }
  }
  // node/triangle/trip/step counters: slots 4, 5, 9, 10 for primary and query
  // launches, 12..15 for the shadow launch
  const int cb = M == kModeShadow ? 12 : 4;
  if (lane_id() == 0) {
    const int slot[6] = {0, 1, 3, 7, 8, 2};
    for (int k = 0; k < 6; ++k)
      if (acc[k]) atomicAdd((unsigned long long*)&counters[slot[k]], (unsigned long long)acc[k]);
    if (COUNT) atomicAdd((unsigned long long*)&counters[M == kModeShadow ? 14 : 9], (unsigned long long)trips);
  }
  if (COUNT) {
    add_counter(counters, cb, c.nodes);
    add_counter(counters, cb + 1, c.tris);
    add_counter(counters, M == kModeShadow ? 15 : 10, c.steps);
    // per-ray step histogram: slot 16/17 = max steps (primary/shadow), 18.. / 25.. = log2 buckets
    const int hb = M == kModeShadow ? 25 : 18;
    for (int k = 0; k < 7; ++k) add_counter(counters, hb + k, hist[k]);
    uint32_t mx = max_steps;
    for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
    if (lane_id() == 0) atomicMax((unsigned long long*)&counters[M == kModeShadow ? 17 : 16], (unsigned long long)mx);
    if (fused) {  // the shadow phase's work goes to the shadow slots
      add_counter(counters, 12, cs.nodes);
      add_counter(counters, 13, cs.tris);
      add_counter(counters, 15, cs.steps);
      for (int k = 0; k < 7; ++k) add_counter(counters, 25 + k, hist_s[k]);
      uint32_t ms = max_steps_s;
      for (int off = 32; off >= 1; off >>= 1) ms = max(ms, (uint32_t)__shfl_xor((int)ms, off, 64));
      if (lane_id() == 0) atomicMax((unsigned long long*)&counters[17], (unsigned long long)ms);
    }
  }
  if (__any(st.overflow) && lane_id() == 0) atomicAdd((unsigned long long*)&counters[6], 1ull);
  if (COUNT && timeline && lane_id() == 0) {
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t* o = timeline + 8 * w;
    o[0] = tl_start, o[1] = tl_exh, o[2] = tl_coop, o[3] = (uint64_t)wall_clock64();
    o[4] = trips, o[5] = trips_coop, o[6] = (uint64_t)__smid(), o[7] = 0;
  }
}

// Completes the hit pixels of the rectangle (j3d/canvas.cpp:788-801, 836-853):
// view-space normal and db_id, and appends one shadow ray per hit pixel to a
// compacted queue.  Each block handles kFinishPerThread * 256 pixels and
// reserves its queue slots with one atomic per pass.
constexpr uint32_t kFinishPerThread = 4;

__global__ void __launch_bounds__(256) k_finish(FrameParams f, const DevObject* __restrict__ objs,
                                                float4* __restrict__ shadow_rays, uint32_t* shadow_count,
                                                uint64_t* counters) {
  __shared__ uint32_t wave_n[4], block_base;
  const uint32_t rw = f.x1 >= f.x0 ? (uint32_t)(f.x1 - f.x0 + 1) : 0u;
  const uint32_t rh = f.y1 >= f.y0 ? (uint32_t)(f.y1 - f.y0 + 1) : 0u;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t total = 0;
  for (uint32_t it = 0; it < kFinishPerThread; ++it) {
    const uint32_t i = (blockIdx.x * kFinishPerThread + it) * blockDim.x + threadIdx.x;
    bool has_shadow = false;
    f4 pos = {0.f, 0.f, 0.f, 0.f}, ld = {0.f, 0.f, 0.f, 0.f};
    uint32_t pix = 0;
    if (i < rw * rh) {
      const uint32_t px = (uint32_t)f.x0 + i % rw, py = (uint32_t)f.y0 + i / rw;
      pix = py * f.w + px;
      char* pp = reinterpret_cast<char*>(f.gbuf) + (size_t)pix * 32;
      uint32_t* pu = reinterpret_cast<uint32_t*>(pp);
      float* pf = reinterpret_cast<float*>(pp);
      const uint32_t tri = pu[4];
      if (tri != kMiss) {
        const uint32_t oi = pu[7];
        const DevObject& ob = objs[oi];
        const float* tnrm = ob.normals + 3 * (size_t)tri;
        f4 n = {gload(tnrm), gload(tnrm + 1), gload(tnrm + 2), 0.f};
        n = mat_vec(f.cs_inv, n);
        n = mat_vec(ob.cs, n);
        bool colored = false;
        const uint32_t word = (f.flags & (J3DG_TEXTURED | J3DG_VERTEXCOLORS))
                                  ? hit_color_word(ob, f.flags, tri, pf[5], pf[6], colored)
                                  : 0u;
        store_hit_word(pp, word, colored);  // mark (j3d/canvas.cpp:801-834)
        pf[1] = n.x;
        pf[2] = n.y;
        pu[7] = ob.db_id;
        if (f.flags & J3DG_SHADOW) {
          const float bu = pf[5], bv = pf[6];
          const uint32_t* tr = ob.idx + 3 * (size_t)tri;
          f4 V[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const float* vv = ob.verts + 3 * (size_t)gload(tr + k);
            V[k] = mat_vec(ob.cs, f4{gload(vv), gload(vv + 1), gload(vv + 2), 1.f});
          }
          float s0 = 1.f - bu - bv;
          pos = f4{V[0].x * s0 + bu * V[1].x + bv * V[2].x, V[0].y * s0 + bu * V[1].y + bv * V[2].y,
                   V[0].z * s0 + bu * V[1].z + bv * V[2].z, V[0].w * s0 + bu * V[1].w + bv * V[2].w};
          if (f.flags & J3DG_DIRECTIONAL_LIGHT)
            ld = f4{f.light_dir[0], f.light_dir[1], f.light_dir[2], 0.f};
          else
            ld = f4{f.light[0] - pos.x, f.light[1] - pos.y, f.light[2] - pos.z, f.light[3] - pos.w};
          has_shadow = true;
        }
      }
    }
    if (!(f.flags & J3DG_SHADOW)) continue;  // uniform
    const uint64_t m = __ballot(has_shadow);
    if (lane_id() == 0) wave_n[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t n = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
      block_base = n ? atomicAdd(shadow_count, n) : 0u;
      total += n;
    }
    __syncthreads();
    uint32_t base = block_base;
    for (uint32_t k = 0; k < wave; ++k) base += wave_n[k];
    if (has_shadow) {
      float4* r = shadow_rays + 3 * (size_t)(base + rank_in_mask(m));
      r[0] = make_float4(pos.x, pos.y, pos.z, pos.w);
      r[1] = make_float4(ld.x, ld.y, ld.z, ld.w);
      r[2] = make_float4(__uint_as_float(pix), 0.f, 0.f, 0.f);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) atomicAdd((unsigned long long*)&counters[2], (unsigned long long)total);
}

// Surface colours of a one-launch frame's hit pixels (j3d/canvas.cpp:803-834):
// the trace left the object index in the db_id word; mark becomes mark | 2
// with r, g, b when the object has a colour (the shadow bit kept), and db_id
// is written.
__global__ void __launch_bounds__(256) k_color(FrameParams f, const DevObject* __restrict__ objs) {
  const uint32_t rw = f.x1 >= f.x0 ? (uint32_t)(f.x1 - f.x0 + 1) : 0u;
  const uint32_t rh = f.y1 >= f.y0 ? (uint32_t)(f.y1 - f.y0 + 1) : 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rw * rh; i += gridDim.x * blockDim.x) {
    const uint32_t pix = ((uint32_t)f.y0 + i / rw) * f.w + (uint32_t)f.x0 + i % rw;
    uint32_t* pu = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(f.gbuf) + (size_t)pix * 32);
    const uint32_t tri = pu[4];
    if (tri == kMiss) continue;
    const uint32_t oi = pu[7] < f.n_obj ? pu[7] : 0u;
    const DevObject& ob = objs[oi];
    const float* pf = reinterpret_cast<const float*>(pu);
    bool colored = false;
    const uint32_t word = hit_color_word(ob, f.flags, tri, pf[5], pf[6], colored);
    if (colored) pu[0] = (pu[0] & 1u) | word;
    pu[7] = ob.db_id;
  }
}

__global__ void k_reset(uint32_t* queue, uint64_t* counters, uint32_t n_counters) {
  uint32_t i = threadIdx.x;
  for (uint32_t k = i; k < kQueueWords; k += blockDim.x) queue[k] = 0u;
  if (i < n_counters) counters[i] = 0ull;
}

// canvas::canvas_to_image (j3d/canvas.cpp:582-670) with the background copy of
// canvas::render_scene (canvas.cpp:895) fused in: one thread per pixel.
__device__ inline void shade_pixel(const ShadeParams& s, uint32_t x, uint32_t y) {
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

__global__ void __launch_bounds__(256) k_shade(ShadeParams s) {
  const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63);
  const uint32_t y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= s.w || y >= s.h) return;
  shade_pixel(s, x, y);
}

// The shading pass of every frame of a batch: blockIdx.z = frame.
__global__ void __launch_bounds__(256) k_shade_batch(const ShadeParams* __restrict__ sps) {
  const ShadeParams s = sps[blockIdx.z];
  const uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63);
  const uint32_t y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= s.w || y >= s.h) return;
  shade_pixel(s, x, y);
}

// Plain streaming copy of known size: the reference point for scaling the
// profiler's HBM byte counters (FETCH_SIZE / WRITE_SIZE).
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
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

// Orders a frame's tiles heaviest first for the next frame with the same
// rectangle (longest-processing-time first: the pixels that take longest run
// while the whole GPU is busy instead of forming the frame's tail).  A tile's
// cost is the loop trips of its 64 pixels; tiles are binned by cost on a log2
// scale (four bins per doubling, 64 bins) and laid out heaviest bin first, each bin in tile order so
// that concurrently running waves stay on neighbouring tiles.  Two small
// launches: per-block bin counts, then k_order_place (each block finds its
// tiles' slots from its band's counts and scatters them stably); frames with
// bands of more than kPlaceMaxBandBlocks blocks scan once (k_order_scan) and
// scatter (k_order_scatter) instead.
constexpr uint32_t kOrderBins = 64;
constexpr uint32_t kOrderBlock = 256;  // tiles per block of k_tile_bins / k_order_scatter
constexpr uint32_t kOrderMinTiles = 8192;  // smallest rectangle (in 8x8 tiles) given a cost order
// A tile holding a pixel of at least this many loop trips goes to the heaviest
// bin whatever its sum, so that the frame's longest pixels start first (28M-
// triangle 1080p frame: 0.784-0.789 ms without; threshold 24: 0.93, 32: 0.83,
// 40: 0.753-0.778, 48: 0.761, 64: 0.777)
constexpr uint32_t kLongPixelTrips = 48;

__device__ inline uint32_t cost_bin(uint32_t c) {
  // 0: up to 2 trips per pixel (background); then four bins per doubling
  if (c < 128u) return 0u;
  const uint32_t b = 1u + (uint32_t)(4.f * __log2f((float)c * (1.f / 128.f)));
  return b < kOrderBins ? b : kOrderBins - 1u;
}

__device__ inline void wave_bin_counts(uint32_t bin, bool valid, uint32_t* counts /* kOrderBins */) {
  for (uint32_t b = 0; b < kOrderBins; ++b) {
    const uint64_t m = __ballot(valid && bin == b);
    if (lane_id() == 0) counts[b] = (uint32_t)__popcll(m);
  }
}

__global__ void __launch_bounds__(kOrderBlock) k_tile_bins(const uint16_t* __restrict__ cost, uint32_t n_tiles,
                                                            uint8_t* __restrict__ tile_bin,
                                                            uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t wc[kOrderBlock / 64][kOrderBins];
  const uint32_t i = blockIdx.x * kOrderBlock + threadIdx.x;
  uint32_t bin = 0;
  if (i < n_tiles) {
    const uint4* p = reinterpret_cast<const uint4*>(cost + (size_t)i * 64);
    uint32_t sum = 0, mx = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint4 q = p[k];
      sum += (q.x & 0xFFFFu) + (q.x >> 16) + (q.y & 0xFFFFu) + (q.y >> 16) + (q.z & 0xFFFFu) + (q.z >> 16) +
             (q.w & 0xFFFFu) + (q.w >> 16);
      if (kLongPixelTrips)
        mx = max(mx, max(max(max(q.x & 0xFFFFu, q.x >> 16), max(q.y & 0xFFFFu, q.y >> 16)),
                         max(max(q.z & 0xFFFFu, q.z >> 16), max(q.w & 0xFFFFu, q.w >> 16))));
    }
    bin = kLongPixelTrips && mx >= kLongPixelTrips ? kOrderBins - 1u : cost_bin(sum);
    tile_bin[i] = (uint8_t)bin;
  }
  wave_bin_counts(bin, i < n_tiles, wc[threadIdx.x >> 6]);
  __syncthreads();
  if (threadIdx.x < kOrderBins) {
    uint32_t c = 0;
    for (uint32_t w = 0; w < kOrderBlock / 64; ++w) c += wc[w][threadIdx.x];
    block_counts[blockIdx.x * kOrderBins + threadIdx.x] = c;
  }
}

// In place: block_counts[block][bin] -> first output slot of (block, bin):
// one exclusive scan over the sequence band 0 (heaviest bin: its blocks in
// order, next bin: ...), band 1, ...  The bands are the k_trace work bands
// (band_first_ray), so every band keeps its own tiles, heaviest first.  One
// block of 1024 threads, each scanning a contiguous segment.
static_assert(kOrderBlock == kBandBlock, "the sort's blocks are the bands' units");
__global__ void __launch_bounds__(1024) k_order_scan(uint32_t n_blocks, uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t part[1024];
  const uint32_t n = n_blocks * kOrderBins, t = threadIdx.x;
  const uint32_t seg = (n + 1023u) / 1024u, lo = min(n, t * seg), hi = min(n, lo + seg);
  auto at = [&](uint32_t k) -> uint32_t& {  // k-th element of the banded heaviest-first sequence
    uint32_t b = 0;
    while (b + 1 < kBands && k >= ((b + 1) * n_blocks / kBands) * kOrderBins) ++b;
    const uint32_t b0 = b * n_blocks / kBands, nbb = (b + 1) * n_blocks / kBands - b0;
    const uint32_t kk = k - b0 * kOrderBins;
    const uint32_t bin = kOrderBins - 1u - kk / nbb, blk = b0 + kk % nbb;
    return block_counts[blk * kOrderBins + bin];
  };
  uint32_t sum = 0;
  for (uint32_t k = lo; k < hi; ++k) sum += at(k);
  part[t] = sum;
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    __syncthreads();
    const uint32_t o = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += o;
  }
  __syncthreads();
  uint32_t run = part[t] - sum;  // exclusive prefix of this segment
  for (uint32_t k = lo; k < hi; ++k) {
    uint32_t& c = at(k);
    const uint32_t v = c;
    c = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kOrderBlock) k_order_scatter(const uint8_t* __restrict__ tile_bin, uint32_t n_tiles,
                                                                const uint32_t* __restrict__ block_first,
                                                                uint32_t* __restrict__ order) {
  __shared__ uint32_t wc[kOrderBlock / 64][kOrderBins];
  const uint32_t i = blockIdx.x * kOrderBlock + threadIdx.x;
  const bool valid = i < n_tiles;
  const uint32_t bin = valid ? tile_bin[i] : 0u;
  const uint32_t w = threadIdx.x >> 6;
  wave_bin_counts(bin, valid, wc[w]);
  uint64_t same = 0;  // lanes of this wave with the same bin
  for (uint32_t b = 0; b < kOrderBins; ++b) {
    const uint64_t mb = __ballot(valid && bin == b);
    if (b == bin) same = mb;
  }
  __syncthreads();
  if (valid) {
    uint32_t pos = block_first[blockIdx.x * kOrderBins + bin];
    for (uint32_t k = 0; k < w; ++k) pos += wc[k][bin];
    order[pos + rank_in_mask(same)] = i;
  }
}

// k_order_scan and k_order_scatter in one launch, for frames whose bands hold
// at most kPlaceMaxBandBlocks blocks: each block sums its band's counts
// itself (64 threads, one bin each) to find where its tiles of each bin go --
// the band's first tile, then the heavier bins of the band, then the same bin
// in the band's earlier blocks -- and scatters its tiles as k_order_scatter.
constexpr uint32_t kPlaceMaxBandBlocks = 512;
static_assert(kOrderBins == 64 && kOrderBlock >= 64, "one wave holds the bins");
__global__ void __launch_bounds__(kOrderBlock) k_order_place(const uint8_t* __restrict__ tile_bin, uint32_t n_tiles,
                                                              const uint32_t* __restrict__ block_counts, uint32_t nb,
                                                              uint32_t* __restrict__ order) {
  __shared__ uint32_t first_pos[kOrderBins];
  __shared__ uint32_t wc[kOrderBlock / 64][kOrderBins];
  const uint32_t blk = blockIdx.x;
  uint32_t b = 0;
  while (b + 1 < kBands && blk >= (b + 1) * nb / kBands) ++b;
  const uint32_t b0 = b * nb / kBands, b1 = (b + 1) * nb / kBands;
  if (threadIdx.x < kOrderBins) {  // wave 0: one bin per lane
    const uint32_t bin = threadIdx.x;
    uint32_t total = 0, before = 0;
    // unrolled: the counts (just written, in L2) are loaded 8 at a time
#pragma unroll 8
    for (uint32_t k = b0; k < b1; ++k) {
      const uint32_t c = block_counts[k * kOrderBins + bin];
      total += c;
      before += k < blk ? c : 0u;
    }
    uint32_t suffix = total;  // the band's tiles in bins >= bin
    for (uint32_t d = 1; d < kOrderBins; d <<= 1) {
      const uint32_t o = (uint32_t)__shfl((int)suffix, (int)min(bin + d, kOrderBins - 1u), 64);
      if (bin + d < kOrderBins) suffix += o;
    }
    first_pos[bin] = b0 * kOrderBlock + (suffix - total) + before;
  }
  const uint32_t i = blk * kOrderBlock + threadIdx.x;
  const bool valid = i < n_tiles;
  const uint32_t bin = valid ? tile_bin[i] : 0u;
  const uint32_t w = threadIdx.x >> 6;
  wave_bin_counts(bin, valid, wc[w]);
  uint64_t same = 0;  // lanes of this wave with the same bin
  for (uint32_t bb = 0; bb < kOrderBins; ++bb) {
    const uint64_t mb = __ballot(valid && bin == bb);
    if (bb == bin) same = mb;
  }
  __syncthreads();
  if (valid) {
    uint32_t pos = first_pos[bin];
    for (uint32_t k = 0; k < w; ++k) pos += wc[k][bin];
    order[pos + rank_in_mask(same)] = i;
  }
}

// A fresh proposal buffer predicts no hit anywhere (plain traversal).
__global__ void k_clear_proposal(float2* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float2(0.f, -1.f);
}

// Working memory of j3dg_trace_rays_device, the ray query that allocates
// nothing: one array per GPU, held until the process ends.  Layout (bytes):
// [0, 1024) queue words, [1024, 2048) statistics counters, [2048, 8192) the
// object table, [8192, end) the traversal-stack spill area.  Host code never
// names `scratch` itself (in host code the name is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress returned for the device.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr size_t kNoAllocCounters = 1024, kNoAllocObjects = 2048, kNoAllocSpill = 8192;
static_assert(kQueueWords * sizeof(uint32_t) <= kNoAllocCounters, "queue words fit their slot");
static_assert(kNoAllocObjects + J3DG_MAX_OBJECTS * sizeof(DevObject) <= kNoAllocSpill, "object table fits its slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];


// ============================================================= host side ===

namespace {

template <int W, int D, int MODE>
void* trace_kernel_wd(bool count) {
  return count ? (void*)&k_trace<W, D, true, MODE> : (void*)&k_trace<W, D, false, MODE>;
}

template <int MODE>
void* trace_kernel_mode(int width, int depth, bool count) {
  if (width == 8) {
    if (depth == 8) return trace_kernel_wd<8, 8, MODE>(count);
    if (depth == 32) return trace_kernel_wd<8, 32, MODE>(count);
    return trace_kernel_wd<8, 16, MODE>(count);
  }
  if (depth == 8) return trace_kernel_wd<4, 8, MODE>(count);
  if (depth == 32) return trace_kernel_wd<4, 32, MODE>(count);
  return trace_kernel_wd<4, 16, MODE>(count);
}

void* trace_kernel(int mode, int width, int depth, bool count) {
  if (mode == (kModeRays | kModeTop)) return trace_kernel_mode<kModeRays | kModeTop>(width, depth, false);
  if (mode == (kModeAnyRays | kModeTop)) return trace_kernel_mode<kModeAnyRays | kModeTop>(width, depth, false);
  if (mode == (kModeFrameColor | kModeTop)) return trace_kernel_mode<kModeFrameColor | kModeTop>(width, depth, count);
  if (mode == (kModeFrame | kModeTop)) return trace_kernel_mode<kModeFrame | kModeTop>(width, depth, count);
  if (mode == kModeBatch) return trace_kernel_mode<kModeBatch>(width, depth, count);
  if (mode == kModeRays) return trace_kernel_mode<kModeRays>(width, depth, false);
  if (mode == kModeAnyRays) return trace_kernel_mode<kModeAnyRays>(width, depth, count);
  if (mode == kModeShadow) return trace_kernel_mode<kModeShadow>(width, depth, count);
  if (mode == kModeFrameColor) return trace_kernel_mode<kModeFrameColor>(width, depth, count);
  return trace_kernel_mode<kModeFrame>(width, depth, count);
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
  int rc = fill_object_table(slots, n, tab.data(), max_stack, width);
  if (rc) return rc;
  if (n > J3DG_MAX_OBJECTS) {
    rc = build_top_level(slots, n);
    if (rc) return rc;
    max_stack += top_depth;  // top-level entries sit below an object's
  }
  return upload_object_table(tab.data(), n);
}

int Ctx::fill_object_table(const uint32_t* slots, uint32_t n, DevObject* tab, uint32_t& max_stack, int& width) {
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
    o.fin = m.d_fin;
    std::memcpy(o.cs, m.cs, sizeof(o.cs));
    std::memcpy(o.inv_cs, m.inv_cs, sizeof(o.inv_cs));
    o.db_id = m.db_id;
    o.width = m.info.width;
    o.vcol = m.d_vcol;
    o.uv = m.d_uv;
    o.tex = m.d_tex;
    o.tex_w = m.tex_w;
    o.tex_h = m.tex_h;
    o.mlp = m.has_proposal ? m.d_mlp_packed : nullptr;
    std::memcpy(o.mlp_norm, m.mlp_norm, sizeof(o.mlp_norm));
    max_stack = std::max(max_stack, m.max_stack);
  }
  return J3DG_OK;
}

// World box of a local box under the column-major pose cs, in double, then
// rounded outward to float.
static void world_box(const float* cs, const float* lo, const float* hi, float* wlo, float* whi) {
  double a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
    for (int k = 0; k < 3; ++k) {
      const double w = (double)cs[k] * p[0] + (double)cs[4 + k] * p[1] + (double)cs[8 + k] * p[2] + (double)cs[12 + k];
      a[k] = std::min(a[k], w);
      b[k] = std::max(b[k], w);
    }
  }
  for (int k = 0; k < 3; ++k) {
    float fl = (float)a[k], fh = (float)b[k];
    if ((double)fl > a[k]) fl = std::nextafter(fl, -INFINITY);
    if ((double)fh < b[k]) fh = std::nextafter(fh, INFINITY);
    wlo[k] = fl;
    whi[k] = fh;
  }
}

// Top level of a scene of more than J3DG_MAX_OBJECTS objects (j3d builds
// qbvh_two_level_with_transformations over its objects every frame,
// canvas.cpp:762): a binary BVH over the objects' world boxes, split at the
// median centroid of the widest axis, one object per leaf.  Rebuilt when an
// object's box or pose changes.
int Ctx::build_top_level(const uint32_t* slots, uint32_t n) {
  if (n > kMaxSceneObjects) return fail(J3DG_ERR_INVALID_ARG, "too many objects in one scene");
  std::vector<float> key;
  key.reserve((size_t)n * 22);
  for (uint32_t i = 0; i < n; ++i) {
    const Mesh& m = meshes[slots[i]];
    key.insert(key.end(), m.info.bb_min, m.info.bb_min + 3);
    key.insert(key.end(), m.info.bb_max, m.info.bb_max + 3);
    key.insert(key.end(), m.cs, m.cs + 16);
  }
  if (d_top && key.size() == top_key.size() &&
      std::memcmp(key.data(), top_key.data(), key.size() * sizeof(float)) == 0)
    return J3DG_OK;
  struct Box {
    float lo[3], hi[3];
    float c[3];
    uint32_t obj;
  };
  std::vector<Box> items(n);
  float radius = 0.f;
  for (uint32_t i = 0; i < n; ++i) {
    const Mesh& m = meshes[slots[i]];
    Box& b = items[i];
    world_box(m.cs, m.info.bb_min, m.info.bb_max, b.lo, b.hi);
    for (int k = 0; k < 3; ++k) {
      b.c[k] = 0.5f * b.lo[k] + 0.5f * b.hi[k];
      radius = std::max({radius, std::fabs(b.lo[k]), std::fabs(b.hi[k]), std::fabs(m.cs[12 + k]),
                         std::fabs(m.info.bb_min[k]), std::fabs(m.info.bb_max[k])});
    }
    b.obj = i;
  }
  if (!std::isfinite(radius)) return fail(J3DG_ERR_INVALID_ARG, "object box or pose is not finite");
  std::vector<TopNode> nodes;
  nodes.reserve(std::max<uint32_t>(n, 2));
  uint32_t depth = 0;
  auto range_box = [&](uint32_t a, uint32_t b, float* out) {
    for (int k = 0; k < 3; ++k) out[k] = INFINITY, out[3 + k] = -INFINITY;
    for (uint32_t i = a; i < b; ++i)
      for (int k = 0; k < 3; ++k) {
        out[k] = std::min(out[k], items[i].lo[k]);
        out[3 + k] = std::max(out[3 + k], items[i].hi[k]);
      }
  };
  // returns the ref of the subtree over items[a, b)
  std::function<uint32_t(uint32_t, uint32_t, uint32_t)> build = [&](uint32_t a, uint32_t b, uint32_t level) {
    if (b - a == 1) return kLeafBit | items[a].obj;
    const uint32_t id = (uint32_t)nodes.size();
    nodes.emplace_back();
    depth = std::max(depth, level);
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = a; i < b; ++i)
      for (int k = 0; k < 3; ++k) clo[k] = std::min(clo[k], items[i].c[k]), chi[k] = std::max(chi[k], items[i].c[k]);
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    const uint32_t mid = a + (b - a) / 2;
    std::nth_element(items.begin() + a, items.begin() + mid, items.begin() + b, [axis](const Box& x, const Box& y) {
      return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.obj < y.obj);
    });
    const uint32_t l = build(a, mid, level + 1);
    const uint32_t r = build(mid, b, level + 1);
    TopNode& nd = nodes[id];
    range_box(a, mid, nd.box[0]);
    range_box(mid, b, nd.box[1]);
    nd.child[0] = l;
    nd.child[1] = r;
    nd.pad[0] = nd.pad[1] = 0;
    return id;
  };
  if (n == 1) {
    TopNode nd;
    range_box(0, 1, nd.box[0]);
    for (int k = 0; k < 6; ++k) nd.box[1][k] = 0.f;
    nd.child[0] = kLeafBit | 0u;
    nd.child[1] = kNone;
    nd.pad[0] = nd.pad[1] = 0;
    nodes.push_back(nd);
    depth = 1;
  } else {
    build(0, n, 1);
  }
  Ctx* ctx = this;
  if (nodes.size() > top_capacity) {
    HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
    if (d_top) hipFree(d_top);
    d_top = nullptr;
    top_capacity = 0;
    HIP_TRY(hipMalloc(&d_top, nodes.size() * sizeof(TopNode)), J3DG_ERR_OUT_OF_MEMORY);
    top_capacity = (uint32_t)nodes.size();
  }
  HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
  HIP_TRY(hipMemcpy(d_top, nodes.data(), nodes.size() * sizeof(TopNode), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  top_depth = depth;
  top_count = (uint32_t)nodes.size();
  top_radius = radius;
  top_key = std::move(key);
  return J3DG_OK;
}

// A hit found in an object's own frame lies within a few float roundings of
// the transforms (relative to the coordinates involved: ray origin, object
// translation and box, distance along the ray) from the world ray; 2^-14 of
// those magnitudes is a wide margin over that.
float Ctx::top_delta(float origin_radius) const {
  return std::ldexp(2.f * (origin_radius + top_radius), -14);
}

int Ctx::upload_object_table(const DevObject* tab, uint32_t n) {
  if (n && n * sizeof(DevObject) == obj_cache.size() &&
      std::memcmp(obj_cache.data(), tab, obj_cache.size()) == 0)
    return J3DG_OK;
  Ctx* ctx = this;
  if (n && n <= J3DG_MAX_OBJECTS) {
    if (!d_obj_ring) {
      HIP_TRY(hipMalloc(&d_obj_ring, (size_t)kObjRing * J3DG_MAX_OBJECTS * sizeof(DevObject)), J3DG_ERR_OUT_OF_MEMORY);
      HIP_TRY(hipHostMalloc(&h_obj_ring, (size_t)kObjRing * J3DG_MAX_OBJECTS * sizeof(DevObject)), J3DG_ERR_OUT_OF_MEMORY);
      for (int i = 0; i < kObjRing; ++i)
        HIP_TRY(hipEventCreateWithFlags(&obj_ring_ev[i], hipEventDisableTiming), J3DG_ERR_LAUNCH);
    }
    // retire the slot in use: every kernel that reads it has been enqueued already
    if (obj_ring_cur >= 0) {
      HIP_TRY(hipEventRecord(obj_ring_ev[obj_ring_cur], stream), J3DG_ERR_LAUNCH);
      obj_ring_used[obj_ring_cur] = true;
    }
    const int j = (obj_ring_cur + 1) % kObjRing;
    if (obj_ring_used[j]) HIP_TRY(hipEventSynchronize(obj_ring_ev[j]), J3DG_ERR_LAUNCH);
    DevObject* h = h_obj_ring + (size_t)j * J3DG_MAX_OBJECTS;
    DevObject* d = d_obj_ring + (size_t)j * J3DG_MAX_OBJECTS;
    std::memcpy(h, tab, n * sizeof(DevObject));
    HIP_TRY(hipMemcpyAsync(d, h, n * sizeof(DevObject), hipMemcpyHostToDevice, stream), J3DG_ERR_LAUNCH);
    obj_ring_cur = j;
    d_objs = d;
  } else if (n) {
    if (n > obj_capacity) {
      HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
      if (d_obj_big) hipFree(d_obj_big);
      d_obj_big = nullptr;
      obj_capacity = 0;
      HIP_TRY(hipMalloc(&d_obj_big, (size_t)n * sizeof(DevObject)), J3DG_ERR_OUT_OF_MEMORY);
      obj_capacity = n;
    }
    HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
    HIP_TRY(hipMemcpy(d_obj_big, tab, n * sizeof(DevObject), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
    d_objs = d_obj_big;
  }
  obj_cache.assign(reinterpret_cast<const uint8_t*>(tab), reinterpret_cast<const uint8_t*>(tab) + n * sizeof(DevObject));
  return J3DG_OK;
}

uint32_t Ctx::trace_grid(uint32_t n_rays, uint32_t D, uint32_t block, uint32_t extra_lds) const {
  const uint32_t waves_per_block = block / 64;
  // stack + object table (or a top level's copy)
  const uint32_t lds = D * block * sizeof(uint2) + J3DG_MAX_OBJECTS * sizeof(DevObject) + extra_lds;
  const uint32_t rays_per_block = 64u * waves_per_block;
  uint32_t blocks_needed = (n_rays + rays_per_block - 1) / rays_per_block;
  // persistent: as many resident blocks as LDS and the 32 waves per CU allow
  uint32_t per_cu = std::max(1u, std::min(160u * 1024u / std::max(lds, 1u), 32u / waves_per_block));
  return std::max(1u, std::min(blocks_needed, (uint32_t)n_cus * per_cu));
}

int Ctx::launch_trace_kernel(int mode, const FrameParams& fp, const RayQuery& q, uint32_t n_rays, uint32_t n_obj,
                             uint32_t max_stack, int width, uint32_t* qcounter, const BatchParams* bpp) {
  Ctx* ctx = this;
  const uint32_t D = tuning.stack_depth;
  const uint32_t block = tuning.block_size;
  const bool top_lds = (mode & kModeTop) && fp.top_count && fp.top_count * sizeof(TopNode) <= kTopLdsBytes;
  const uint32_t top_bytes = top_lds ? fp.top_count * (uint32_t)sizeof(TopNode) : 0u;
  const uint32_t lds = D * block * sizeof(uint2) + top_bytes;
  uint32_t grid = trace_grid(n_rays, D, block, top_bytes);
  uint64_t lanes = (uint64_t)grid * block;
  uint32_t cap = max_stack > D ? max_stack - D + 1 : 1;
  int rc = ensure_spill(lanes, cap);
  if (rc) return rc;
  void* fn = trace_kernel(mode, width, (int)D, counting);
  uint2* sp = spill;
  uint32_t* qu = qcounter;
  uint64_t* cnt = d_counters;
  FrameParams f = fp;
  f.top_lds = top_lds ? 1u : 0u;
  RayQuery rq = q;
  uint32_t nr = n_rays, no = n_obj;
  uint32_t batch = 64u * tuning.rays_per_item;
  uint32_t refill = tuning.refill_min;
  uint32_t coop = tuning.coop_tail;
  // a batch's frames index their objects in the batch's own table
  const DevObject* objs = mode == kModeBatch ? d_bobjs : d_objs;
  uint64_t* tl = nullptr;
  if (counting && (mode == kModeFrame || mode == kModeFrameColor)) {
    uint64_t waves = lanes / 64;
    if (waves * 8 > timeline_capacity) {
      if (d_timeline) hipFree(d_timeline);
      d_timeline = nullptr;
      timeline_capacity = 0;
      HIP_TRY(hipMalloc(&d_timeline, waves * 8 * sizeof(uint64_t)), J3DG_ERR_OUT_OF_MEMORY);
      timeline_capacity = waves * 8;
    }
    tl = d_timeline;
    timeline_waves = (uint32_t)waves;
  }
  BatchParams bp;
  std::memset(&bp, 0, sizeof(bp));
  if (bpp) bp = *bpp;
  void* args[] = {&f, &rq, &nr, &objs, &no, &sp, &cap, &qu, &cnt, &batch, &refill, &coop, &tl, &bp};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int Ctx::ensure_shadow_queue(uint64_t rays) {
  if (rays <= shadow_capacity) return J3DG_OK;
  Ctx* ctx = this;
  HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
  if (d_shadow) hipFree(d_shadow);
  d_shadow = nullptr;
  shadow_capacity = 0;
  HIP_TRY(hipMalloc(&d_shadow, rays * 3 * sizeof(float4)), J3DG_ERR_OUT_OF_MEMORY);
  shadow_capacity = rays;
  return J3DG_OK;
}

// The trace stage.  Fused (default): one k_trace launch whose lanes complete
// their hits and trace the shadow rays.  Split: primary trace, k_finish
// (normals, db_id, shadow queue) and the shadow trace.
int Ctx::launch_trace(const FrameParams& fp_in, uint32_t max_stack, int width, bool learned, Target* tg) {
  (void)learned;
  Ctx* ctx = this;
  RayQuery q;
  std::memset(&q, 0, sizeof(q));
  hipEvent_t* e = cur_ev();
  FrameParams fp = fp_in;
  // scenes with a top level always run as one launch (the split pipeline has no top-level variant)
  const int top = fp.top_nodes ? kModeTop : 0;
  fp.fused = ((tuning.split_passes == 0 || top) && fp.n_obj) ? 1u : 0u;
  const uint32_t rect_px = fp.n_tiles * 64u;
  // the cost order pays from ~8192 tiles up: at 640x480 (4800 tiles) its three
  // sorting launches cost 20 us more than they save, at 1280x720 (14400) they save 24 us
  const bool order = tg && tuning.cost_order && fp.n_obj && fp.n_tiles >= kOrderMinTiles;
  if (order) {
    if (fp.n_tiles > tg->order_capacity) {
      HIP_TRY(hipStreamSynchronize(stream), J3DG_ERR_LAUNCH);
      if (tg->pixel_cost) hipFree(tg->pixel_cost);
      if (tg->tile_cost) hipFree(tg->tile_cost);
      if (tg->tile_order) hipFree(tg->tile_order);
      tg->pixel_cost = nullptr, tg->tile_cost = nullptr, tg->tile_order = nullptr;
      tg->order_capacity = 0;
      tg->order_valid = false;
      HIP_TRY(hipMalloc(&tg->pixel_cost, (size_t)fp.n_tiles * 64 * sizeof(uint16_t)), J3DG_ERR_OUT_OF_MEMORY);
      const size_t nb = (fp.n_tiles + kOrderBlock - 1) / kOrderBlock;
      HIP_TRY(hipMalloc(&tg->tile_cost, ((fp.n_tiles + 3) / 4 + nb * kOrderBins) * sizeof(uint32_t)),
              J3DG_ERR_OUT_OF_MEMORY);
      HIP_TRY(hipMalloc(&tg->tile_order, (size_t)fp.n_tiles * sizeof(uint32_t)), J3DG_ERR_OUT_OF_MEMORY);
      tg->order_capacity = fp.n_tiles;
    }
    const bool same = tg->order_valid && tg->order_rect[0] == fp.x0 && tg->order_rect[1] == fp.y0 &&
                      tg->order_rect[2] == fp.x1 && tg->order_rect[3] == fp.y1;
    fp.tile_order = same ? tg->tile_order : nullptr;
    fp.pixel_cost = tg->pixel_cost;
  }
  HIP_TRY(hipEventRecord(e[0], stream), J3DG_ERR_LAUNCH);
  const bool color = fp.fused && (fp.flags & (J3DG_TEXTURED | J3DG_VERTEXCOLORS));
  int rc = launch_trace_kernel((color ? kModeFrameColor : kModeFrame) | top, fp, q, rect_px, fp.n_obj, max_stack,
                               width, d_queue);
  if (rc) return rc;
  if (color) {
    const uint32_t n = (uint32_t)std::min<uint64_t>(((uint64_t)fp.n_tiles * 64 + 255) / 256, (uint64_t)n_cus * 16);
    hipLaunchKernelGGL(k_color, dim3(std::max(1u, n)), dim3(256), 0, stream, fp, (const DevObject*)d_objs);
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  }
  if (order) {
    // tile_cost holds the per-tile bins (bytes) followed by the per-block counts
    const uint32_t nb = (fp.n_tiles + kOrderBlock - 1) / kOrderBlock;
    uint8_t* bins = reinterpret_cast<uint8_t*>(tg->tile_cost);
    uint32_t* counts = tg->tile_cost + (fp.n_tiles + 3) / 4;
    hipLaunchKernelGGL(k_tile_bins, dim3(nb), dim3(kOrderBlock), 0, stream, (const uint16_t*)tg->pixel_cost,
                       fp.n_tiles, bins, counts);
    if ((nb + kBands - 1) / kBands <= kPlaceMaxBandBlocks) {
      hipLaunchKernelGGL(k_order_place, dim3(nb), dim3(kOrderBlock), 0, stream, (const uint8_t*)bins, fp.n_tiles,
                         (const uint32_t*)counts, nb, tg->tile_order);
    } else {
      hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, stream, nb, counts);
      hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(kOrderBlock), 0, stream, (const uint8_t*)bins, fp.n_tiles,
                         (const uint32_t*)counts, tg->tile_order);
    }
    HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    tg->order_valid = true;
    tg->order_rect[0] = fp.x0, tg->order_rect[1] = fp.y0, tg->order_rect[2] = fp.x1, tg->order_rect[3] = fp.y1;
  }
  HIP_TRY(hipEventRecord(e[1], stream), J3DG_ERR_LAUNCH);
  HIP_TRY(hipEventRecord(e[6], stream), J3DG_ERR_LAUNCH);
  if (fp.n_obj && !fp.fused) {
    const bool shadow = (fp.flags & J3DG_SHADOW) != 0;
    if (shadow) {
      rc = ensure_shadow_queue(rect_px);
      if (rc) return rc;
    }
    uint32_t rw = fp.x1 >= fp.x0 ? (uint32_t)(fp.x1 - fp.x0 + 1) : 0;
    uint32_t rh = fp.y1 >= fp.y0 ? (uint32_t)(fp.y1 - fp.y0 + 1) : 0;
    uint32_t n = rw * rh;
    if (n) {
      const uint32_t per_block = 256u * kFinishPerThread;
      hipLaunchKernelGGL(k_finish, dim3((n + per_block - 1) / per_block), dim3(256), 0, stream, fp,
                         (const DevObject*)d_objs, d_shadow, d_queue + 1, d_counters);
      HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
    }
    if (shadow && n) {
      q.rays = reinterpret_cast<const float*>(d_shadow);
      q.count = d_queue + 1;
      rc = launch_trace_kernel(kModeShadow, fp, q, n, fp.n_obj, max_stack, width, d_queue + 2);
      if (rc) return rc;
    }
  }
  HIP_TRY(hipEventRecord(e[7], stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int Ctx::reset_counters() {
  Ctx* ctx = this;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, d_queue, d_counters, 32u);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int Ctx::launch_shade(const ShadeParams& sp) {
  Ctx* ctx = this;
  dim3 grid((sp.w + 63) / 64, (sp.h + 3) / 4);
  hipEvent_t* e = cur_ev();
  HIP_TRY(hipEventRecord(e[2], stream), J3DG_ERR_LAUNCH);
  hipLaunchKernelGGL(k_shade, grid, dim3(256), 0, stream, sp);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_TRY(hipEventRecord(e[3], stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // namespace j3d

// ================================================================== C ABI ===

using j3d::Ctx;
using j3d::Mesh;
using j3d::Target;


namespace j3d {

// Mesh bookkeeping shared by both builders: pose, bounding box (scene.cpp:22,
// mesh.cpp:39-58) and a free slot.
static void init_mesh(Mesh& m, uint32_t db_id, const float* cs, const float* xyz, uint64_t n_vertices, uint64_t n_triangles) {
  m.live = true;
  m.db_id = db_id;
  if (cs)
    std::memcpy(m.cs, cs, sizeof(m.cs));
  else
    for (int i = 0; i < 16; ++i) m.cs[i] = (i % 5 == 0) ? 1.f : 0.f;
  invert_orthonormal(m.cs, m.inv_cs);
  m.info.n_vertices = n_vertices;
  m.info.n_triangles = n_triangles;
  if (!n_vertices) return;
  // std::min / std::max folded over the vertices from the first one.  Threads
  // fold chunks from +-inf and are combined in chunk order: the same result
  // for every input (a NaN is skipped by both folds, and the first of equal
  // values is kept), except a NaN in the first vertex, which poisons its axis
  // in the sequential fold and is applied explicitly.
  const uint64_t per = 1ull << 20;
  const uint64_t chunks = (n_vertices + per - 1) / per;
  std::vector<float> part(6 * chunks);
  auto fold = [&](uint64_t c) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint64_t a = c * per, b = std::min(n_vertices, a + per);
    for (uint64_t i = a; i < b; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], xyz[3 * i + k]);
        hi[k] = std::max(hi[k], xyz[3 * i + k]);
      }
    for (int k = 0; k < 3; ++k) part[6 * c + k] = lo[k], part[6 * c + 3 + k] = hi[k];
  };
  const uint64_t threads = std::min<uint64_t>(chunks, std::max(1u, std::thread::hardware_concurrency()));
  if (threads <= 1) {
    for (uint64_t c = 0; c < chunks; ++c) fold(c);
  } else {
    std::vector<std::thread> pool;
    for (uint64_t t = 0; t < threads; ++t)
      pool.emplace_back([&, t]() {
        for (uint64_t c = t; c < chunks; c += threads) fold(c);
      });
    for (auto& th : pool) th.join();
  }
  for (int k = 0; k < 3; ++k) {
    float lo = xyz[k], hi = xyz[k];
    for (uint64_t c = 0; c < chunks; ++c) {
      lo = std::min(lo, part[6 * c + k]);
      hi = std::max(hi, part[6 * c + 3 + k]);
    }
    m.info.bb_min[k] = lo;
    m.info.bb_max[k] = hi;
  }
}

static uint32_t insert_mesh(Ctx* ctx, Mesh&& m) {
  uint32_t s = (uint32_t)ctx->meshes.size();
  for (uint32_t i = 0; i < ctx->meshes.size(); ++i)
    if (!ctx->meshes[i].live) {
      s = i;
      break;
    }
  if (s == ctx->meshes.size()) ctx->meshes.emplace_back();
  ctx->meshes[s] = std::move(m);
  ctx->obj_cache.clear();
  return s;
}

// J3DG_BUILD_DEVICE: raw arrays up, normals, completion records and the
// linear BVH built on the GPU (bvh_gpu.hip); the normals stay on the device
// (j3dg_mesh_normals reads them back).
static int upload_mesh_device(Ctx* ctx, uint32_t db_id, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                       uint64_t n_triangles, const float* cs, const BuildOptions& bo, uint32_t* slot) {
  auto t0 = std::chrono::steady_clock::now();
  Mesh m;
  init_mesh(m, db_id, cs, xyz, n_vertices, n_triangles);
  double upload_s = 0.0;
  int rc = device_build_mesh(ctx, m, xyz, n_vertices, tris, n_triangles, bo.width, bo.leaf_size, &upload_s);
  if (rc != J3DG_OK) {
    m.release();
    return rc;
  }
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  m.info.upload_seconds = upload_s;
  m.info.build_seconds = total - upload_s;
  *slot = insert_mesh(ctx, std::move(m));
  return J3DG_OK;
}

}  // namespace j3d

extern "C" {

int j3dg_abi_version(void) { return J3DG_ABI_VERSION; }

// Device address of `scratch` on each GPU, looked up once per device (with
// that device current) by the first j3dg_create on it.
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;

static int noalloc_scratch_setup(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(j3d::scratch)) != hipSuccess || !p) return J3DG_ERR_LAUNCH;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return J3DG_OK;
}

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
  for (auto& e : ctx->splat_ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(J3DG_ERR_LAUNCH);
  ctx->ring.assign(8 * Ctx::kRing, nullptr);
  for (auto& e : ctx->ring)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(J3DG_ERR_LAUNCH);
  if (hipMalloc(&ctx->d_queue, kQueueWords * sizeof(uint32_t)) != hipSuccess) return cleanup(J3DG_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->d_counters, 32 * sizeof(uint64_t)) != hipSuccess) return cleanup(J3DG_ERR_OUT_OF_MEMORY);
  if (noalloc_scratch_setup(device, &ctx->noalloc_scratch) != J3DG_OK) return cleanup(J3DG_ERR_LAUNCH);
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
  for (auto& c : ctx->clouds) c.release();
  ctx->release_buffers();
  for (auto& e : ctx->ev)
    if (e) hipEventDestroy(e);
  for (auto& e : ctx->splat_ev)
    if (e) hipEventDestroy(e);
  for (auto& e : ctx->ring)
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
  if (n.refill_min == 0) n.refill_min = 32;
  if (n.refill_min > 64) return ctx->fail(J3DG_ERR_INVALID_ARG, "refill_min must be in [1, 64]");
  if (n.stack_depth != 8 && n.stack_depth != 16 && n.stack_depth != 32)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "stack_depth must be 8, 16 or 32");
  if (n.rays_per_item != 1 && n.rays_per_item != 2 && n.rays_per_item != 4)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "rays_per_item must be 1, 2 or 4");
  if (n.block_size != 64 && n.block_size != 128 && n.block_size != 256)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "block_size must be 64, 128 or 256");
  if (n.split_passes > 1) return ctx->fail(J3DG_ERR_INVALID_ARG, "split_passes must be 0 or 1");
  if (n.cost_order > 1) return ctx->fail(J3DG_ERR_INVALID_ARG, "cost_order must be 0 or 1");
  if (n.coop_tail > 4) return ctx->fail(J3DG_ERR_INVALID_ARG, "coop_tail must be 0 .. 4");
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
  uint64_t c[32];
  HIP_CALL(hipMemcpy(c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  std::memset(st, 0, sizeof(*st));
  st->shadow_node_visits = c[12];
  st->shadow_tri_tests = c[13];
  st->shadow_wave_trips = c[14];
  st->shadow_lane_steps = c[15];
  st->max_ray_steps = c[16];
  st->max_shadow_ray_steps = c[17];
  for (int k = 0; k < 7; ++k) {
    st->ray_steps_log2[k] = c[18 + k];
    st->shadow_ray_steps_log2[k] = c[25 + k];
  }
  st->primary_rays = c[0];
  st->hits = c[1];
  st->shadow_rays = c[2];
  st->shadowed = c[3];
  st->node_visits = c[4];
  st->tri_tests = c[5];
  st->stack_overflow = c[6];
  st->learned_resolved = c[7];
  st->learned_fallback = c[8];
  st->wave_trips = c[9];
  st->lane_steps = c[10];
  float ms = 0.f;
  const hipEvent_t* e = ctx->last_ev ? ctx->last_ev : ctx->ev;
  if (ctx->traced && hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess) st->trace_ms = ms;
  if (ctx->shaded && hipEventElapsedTime(&ms, e[2], e[3]) == hipSuccess) st->shade_ms = ms;
  if (ctx->mlp_ran && hipEventElapsedTime(&ms, e[4], e[5]) == hipSuccess) st->mlp_ms = ms;
  if (ctx->traced && hipEventElapsedTime(&ms, e[6], e[7]) == hipSuccess) st->shadow_ms = ms;
  st->splat_ms = ctx->splat_ms;
  return J3DG_OK;
}

int j3dg_reset_timings(j3dg_ctx* h) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  ctx->ring_n = 0;
  return J3DG_OK;
}

int j3dg_get_timings(j3dg_ctx* h, float* trace_ms, float* shade_ms, float* mlp_ms, float* shadow_ms, uint32_t max,
                     uint32_t* n) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!n) return ctx->fail(J3DG_ERR_INVALID_ARG, "null count");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  uint32_t cnt = std::min(ctx->ring_n, Ctx::kRing);
  uint32_t first = ctx->ring_n - cnt;
  uint32_t k = 0;
  for (uint32_t i = first; i < ctx->ring_n && k < max; ++i, ++k) {
    const hipEvent_t* e = &ctx->ring[8 * (i % Ctx::kRing)];
    uint32_t fl = ctx->ring_flags[i % Ctx::kRing];
    float ms = 0.f;
    if (trace_ms) trace_ms[k] = ((fl & 1) && hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess) ? ms : 0.f;
    if (shade_ms) shade_ms[k] = ((fl & 2) && hipEventElapsedTime(&ms, e[2], e[3]) == hipSuccess) ? ms : 0.f;
    if (mlp_ms) mlp_ms[k] = ((fl & 4) && hipEventElapsedTime(&ms, e[4], e[5]) == hipSuccess) ? ms : 0.f;
    if (shadow_ms) shadow_ms[k] = ((fl & 1) && hipEventElapsedTime(&ms, e[6], e[7]) == hipSuccess) ? ms : 0.f;
  }
  *n = k;
  return J3DG_OK;
}

int j3dg_upload_mesh(j3dg_ctx* h, uint32_t db_id, const float* xyz, uint64_t n_vertices, const uint32_t* tris,
                     uint64_t n_triangles, const float cs[16], const j3dg_build_opts* opts, uint32_t* slot) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!slot || (!xyz && n_vertices) || (!tris && n_triangles))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "null mesh array");
  if (cs)
    for (int i = 0; i < 16; ++i)
      if (!std::isfinite(cs[i])) return ctx->fail(J3DG_ERR_INVALID_ARG, "object pose is not finite");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  auto t0 = std::chrono::steady_clock::now();
  BuildOptions bo;
  bo.width = 8;
  if (opts) {
    if (opts->width) bo.width = (int)opts->width;
    bo.threads = (int)opts->threads;
  }
  // default leaf size: two triangles for 8-wide trees (a leaf is then one
  // 96-byte fetch, like a node), four for 4-wide ones
  bo.leaf_size = bo.width == 8 ? 2 : 4;
  if (opts && opts->leaf_size) bo.leaf_size = (int)opts->leaf_size;
  const uint32_t builder = opts ? opts->builder : J3DG_BUILD_HOST;
  if (builder != J3DG_BUILD_HOST && builder != J3DG_BUILD_DEVICE)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "unknown BVH builder");
  if (builder == J3DG_BUILD_DEVICE && n_triangles >= 2)
    return j3d::upload_mesh_device(ctx, db_id, xyz, n_vertices, tris, n_triangles, cs, bo, slot);
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
  j3d::init_mesh(m, db_id, cs, xyz, n_vertices, n_triangles);
  m.info.n_nodes = bvh.n_nodes;
  m.info.width = (uint32_t)bvh.width;
  m.info.max_depth = bvh.max_depth;
  m.max_stack = bvh.max_stack + 1;
  auto t1 = std::chrono::steady_clock::now();
  m.info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
  m.normals_host = std::move(normals);
  size_t nb = bvh.nodes.size() * sizeof(uint32_t), tb = bvh.tris.size() * sizeof(float);
  size_t vb = 3 * n_vertices * sizeof(float), ib = 3 * n_triangles * sizeof(uint32_t);
  size_t mb = 3 * n_triangles * sizeof(float);
  size_t fb = 3 * n_triangles * sizeof(float4);
  const size_t fb_alloc = fb + 3 * sizeof(float4);  // a spare record, as for the triangles
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)); };
  // one spare triangle record: the traversal fetches 96 bytes at a leaf, so a
  // one-triangle leaf at the end of the array must still read in bounds
  const size_t tb_alloc = tb + 3 * sizeof(float4);
  if (alloc((void**)&m.d_nodes, nb) != hipSuccess || alloc((void**)&m.d_tris, tb_alloc) != hipSuccess ||
      alloc((void**)&m.d_verts, vb) != hipSuccess || alloc((void**)&m.d_idx, ib) != hipSuccess ||
      alloc((void**)&m.d_normals, mb) != hipSuccess || alloc((void**)&m.d_fin, fb_alloc) != hipSuccess) {
    m.release();
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device allocation for mesh failed");
  }
  bool ok = hipMemcpy(m.d_nodes, bvh.nodes.data(), nb, hipMemcpyHostToDevice) == hipSuccess &&
            (!tb || hipMemcpy(m.d_tris, bvh.tris.data(), tb, hipMemcpyHostToDevice) == hipSuccess) &&
            hipMemset(reinterpret_cast<char*>(m.d_tris) + tb, 0, tb_alloc - tb) == hipSuccess &&
            hipMemset(reinterpret_cast<char*>(m.d_fin) + fb, 0, fb_alloc - fb) == hipSuccess &&
            (!vb || hipMemcpy(m.d_verts, xyz, vb, hipMemcpyHostToDevice) == hipSuccess) &&
            (!ib || hipMemcpy(m.d_idx, tris, ib, hipMemcpyHostToDevice) == hipSuccess) &&
            (!mb || hipMemcpy(m.d_normals, m.normals_host.data(), mb, hipMemcpyHostToDevice) == hipSuccess);
  // hit completion records, by triangle id: (v0, n.x), (v1, n.y), (v2, n.z), the
  // same floats the G-buffer normal and the shadow origin are computed from
  // (j3d/canvas.cpp:790-792, 844-847), packed so a hit needs one round trip
  if (ok && n_triangles) {
    const uint64_t chunk = std::min<uint64_t>(n_triangles, 1ull << 22);
    std::vector<float4> buf(3 * (size_t)chunk);
    const float* nh = m.normals_host.data();
    const int threads = bo.threads > 0 ? bo.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    for (uint64_t c0 = 0; c0 < n_triangles && ok; c0 += chunk) {
      const uint64_t c1 = std::min<uint64_t>(n_triangles, c0 + chunk);
      const uint64_t per = (c1 - c0 + threads - 1) / threads;
      std::vector<std::thread> pool;
      for (int t = 0; t < threads; ++t) {
        const uint64_t a = c0 + std::min<uint64_t>(c1 - c0, t * per), b = c0 + std::min<uint64_t>(c1 - c0, (t + 1) * per);
        if (a < b)
          pool.emplace_back([&, a, b]() {
            for (uint64_t i = a; i < b; ++i)
              for (int k = 0; k < 3; ++k) {
                const float* v = xyz + 3 * (size_t)tris[3 * i + k];
                buf[3 * (size_t)(i - c0) + k] = make_float4(v[0], v[1], v[2], nh[3 * i + k]);
              }
          });
      }
      for (auto& th : pool) th.join();
      ok = hipMemcpy(m.d_fin + 3 * (size_t)c0, buf.data(), 3 * (c1 - c0) * sizeof(float4), hipMemcpyHostToDevice) ==
           hipSuccess;
    }
  }
  if (!ok) {
    m.release();
    return ctx->fail(J3DG_ERR_LAUNCH, "mesh upload copy failed");
  }
  auto t2 = std::chrono::steady_clock::now();
  m.info.upload_seconds = std::chrono::duration<double>(t2 - t1).count();
  m.info.device_bytes = nb + tb + vb + ib + mb + fb;
  *slot = j3d::insert_mesh(ctx, std::move(m));
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

int j3dg_set_mesh_colors(j3dg_ctx* h, uint32_t slot, const float* vertex_colors, const float* uv,
                         const uint32_t* texture, uint32_t tex_w, uint32_t tex_h) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  if (uv && (!texture || tex_w == 0 || tex_h == 0))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "uv coordinates need a texture of nonzero size");
  if (!uv && texture) return ctx->fail(J3DG_ERR_INVALID_ARG, "a texture needs uv coordinates");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  Mesh& m = ctx->meshes[slot];
  if (m.d_vcol) hipFree(m.d_vcol);
  if (m.d_uv) hipFree(m.d_uv);
  if (m.d_tex) hipFree(m.d_tex);
  m.d_vcol = nullptr, m.d_uv = nullptr, m.d_tex = nullptr;
  m.tex_w = m.tex_h = 0;
  ctx->obj_cache.clear();
  ctx->invalidate_graphs();
  const size_t cb = 3 * (size_t)m.info.n_vertices * sizeof(float);
  const size_t ub = 6 * (size_t)m.info.n_triangles * sizeof(float);
  const size_t xb = (size_t)tex_w * tex_h * sizeof(uint32_t);
  bool ok = true;
  if (vertex_colors && cb)
    ok = hipMalloc(&m.d_vcol, cb) == hipSuccess && hipMemcpy(m.d_vcol, vertex_colors, cb, hipMemcpyHostToDevice) == hipSuccess;
  if (ok && uv && ub)
    ok = hipMalloc(&m.d_uv, ub) == hipSuccess && hipMemcpy(m.d_uv, uv, ub, hipMemcpyHostToDevice) == hipSuccess &&
         hipMalloc(&m.d_tex, xb) == hipSuccess && hipMemcpy(m.d_tex, texture, xb, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    if (m.d_vcol) hipFree(m.d_vcol);
    if (m.d_uv) hipFree(m.d_uv);
    if (m.d_tex) hipFree(m.d_tex);
    m.d_vcol = nullptr, m.d_uv = nullptr, m.d_tex = nullptr;
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "device allocation or copy for mesh colours failed");
  }
  if (m.d_tex) m.tex_w = tex_w, m.tex_h = tex_h;
  return J3DG_OK;
}

int j3dg_set_mesh_cs(j3dg_ctx* h, uint32_t slot, const float cs[16]) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!cs) return ctx->fail(J3DG_ERR_INVALID_ARG, "null matrix");
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(cs[i])) return ctx->fail(J3DG_ERR_INVALID_ARG, "object pose is not finite");
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
  const Mesh& m = ctx->meshes[slot];
  const auto& n = m.normals_host;
  if (!n.empty()) {
    std::memcpy(out, n.data(), n.size() * sizeof(float));
  } else if (m.info.n_triangles && m.d_normals) {
    HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
    HIP_CALL(hipMemcpy(out, m.d_normals, 3 * m.info.n_triangles * sizeof(float), hipMemcpyDeviceToHost),
             J3DG_ERR_LAUNCH);
  }
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

int j3dg_read_target_strided(j3dg_ctx* h, uint32_t target, j3dg_pixel* gbuf, uint32_t gbuf_stride, uint32_t* rgba,
                             uint32_t rgba_stride) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& t = ctx->targets[target];
  if ((gbuf && gbuf_stride < t.w) || (rgba && rgba_stride < t.w))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "row stride below the target width");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  // one 2-D copy per buffer: rows of w elements from the packed device image into the strided host one
  if (gbuf)
    HIP_CALL(hipMemcpy2D(gbuf, (size_t)gbuf_stride * sizeof(j3dg_pixel), t.gbuf, (size_t)t.w * sizeof(j3dg_pixel),
                         (size_t)t.w * sizeof(j3dg_pixel), t.h, hipMemcpyDeviceToHost),
             J3DG_ERR_LAUNCH);
  if (rgba)
    HIP_CALL(hipMemcpy2D(rgba, (size_t)rgba_stride * sizeof(uint32_t), t.rgba, (size_t)t.w * sizeof(uint32_t),
                         (size_t)t.w * sizeof(uint32_t), t.h, hipMemcpyDeviceToHost),
             J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

static int ensure_proposal(Ctx* ctx, Target& tg) {
  if (tg.proposal) return J3DG_OK;
  HIP_CALL(hipMalloc(&tg.proposal, (size_t)tg.w * tg.h * sizeof(float2)), J3DG_ERR_OUT_OF_MEMORY);
  hipLaunchKernelGGL(k_clear_proposal, dim3((unsigned)(((size_t)tg.w * tg.h + 255) / 256)), dim3(256), 0, ctx->stream,
                     tg.proposal, (uint64_t)tg.w * tg.h);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_write_target_proposal(j3dg_ctx* h, uint32_t target, const float* pairs) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!pairs) return ctx->fail(J3DG_ERR_INVALID_ARG, "null proposal");
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& tg = ctx->targets[target];
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  int rc = ensure_proposal(ctx, tg);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(tg.proposal, pairs, (size_t)tg.w * tg.h * sizeof(float2), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_read_target_proposal(j3dg_ctx* h, uint32_t target, float* pairs) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!pairs) return ctx->fail(J3DG_ERR_INVALID_ARG, "null proposal");
  if (target >= ctx->targets.size() || !ctx->targets[target].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target");
  Target& tg = ctx->targets[target];
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  int rc = ensure_proposal(ctx, tg);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(pairs, tg.proposal, (size_t)tg.w * tg.h * sizeof(float2), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
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
  // the scene's mesh slots: objects[] (up to J3DG_MAX_OBJECTS), or object_list (any number)
  const uint32_t* slots = fd->object_list ? fd->object_list : fd->objects;
  if (!fd->object_list && fd->n_objects > J3DG_MAX_OBJECTS)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "more than J3DG_MAX_OBJECTS objects need object_list");
  if (fd->n_objects > j3d::kMaxSceneObjects) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many objects");
  if ((fd->flags & J3DG_LEARNED) && fd->n_objects > J3DG_MAX_OBJECTS && !(fd->flags & J3DG_NO_SCENE))
    return ctx->fail(J3DG_ERR_UNSUPPORTED, "learned frames take at most J3DG_MAX_OBJECTS objects");
  if ((stages & J3DG_STAGE_SHADE) && !ctx->d_matcap) return ctx->fail(J3DG_ERR_INVALID_ARG, "no matcap set");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  ctx->traced = ctx->shaded = ctx->mlp_ran = false;
  const uint64_t npx = (uint64_t)tg.w * tg.h;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->d_queue, ctx->d_counters, 32u);
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
        rc = ctx->build_object_table(slots, fd->n_objects, max_stack, width);
        if (rc) return rc;
        if (fd->n_objects > J3DG_MAX_OBJECTS) {
          // primary rays start at the camera, shadow rays at hits inside the scene's boxes
          const float cam = std::max({std::fabs(fp.cs[12]), std::fabs(fp.cs[13]), std::fabs(fp.cs[14])});
          fp.top_nodes = ctx->d_top;
          fp.top_delta = ctx->top_delta(cam);
          fp.top_count = ctx->top_count;
        }
      }
      if ((fd->flags & J3DG_LEARNED) && fp.n_obj) {
        bool any_net = false;
        for (uint32_t i = 0; i < fd->n_objects; ++i) any_net = any_net || ctx->meshes[slots[i]].has_proposal;
        const bool keep = (fd->flags & J3DG_KEEP_PROPOSAL) != 0;
        if (keep || any_net) {
          rc = ensure_proposal(ctx, tg);
          if (rc) return rc;
          fp.proposal = tg.proposal;
          if (!keep) {
            rc = ctx->launch_mlp_frame(fp, fp.n_obj);
            if (rc) return rc;
          }
        }
      }
      rc = ctx->launch_trace(fp, std::max<uint32_t>(max_stack, 1), width, fp.proposal != nullptr, &tg);
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
  {
    uint32_t fl = (ctx->traced ? 1u : 0u) | (ctx->shaded ? 2u : 0u) | (ctx->mlp_ran ? 4u : 0u);
    ctx->last_ev = ctx->cur_ev();
    ctx->ring_flags[ctx->ring_n % Ctx::kRing] = fl;
    ctx->ring_n++;
  }
  if (gbuf_out) HIP_CALL(hipMemcpyAsync(gbuf_out, tg.gbuf, npx * sizeof(j3dg_pixel), hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
  if (rgba_out) HIP_CALL(hipMemcpyAsync(rgba_out, tg.rgba, npx * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream), J3DG_ERR_LAUNCH);
  if (gbuf_out || rgba_out) HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

static int trace_rays_impl(Ctx* ctx, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                           uint32_t* obj, uint32_t* tri, float* t, float* u, float* v, uint8_t* hit, bool any) {
  if (!rays && n) return ctx->fail(J3DG_ERR_INVALID_ARG, "null rays");
  if (n_slots > j3d::kMaxSceneObjects || (n_slots && !slots)) return ctx->fail(J3DG_ERR_INVALID_ARG, "bad object list");
  if (n == 0) return J3DG_OK;
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  uint32_t max_stack = 1;
  int width = 4;
  if (n_slots) {
    int rc = ctx->build_object_table(slots, n_slots, max_stack, width);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->d_queue, ctx->d_counters, 32u);
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
  if (n > 0xFFFFFFFFull / 2) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many rays in one call");
  uint32_t* o_obj = reinterpret_cast<uint32_t*>(d_out);
  uint32_t* o_tri = o_obj + n;
  float* o_t = reinterpret_cast<float*>(o_tri + n);
  float* o_u = o_t + n;
  float* o_v = o_u + n;
  RayQuery q;
  q.rays = d_rays;
  q.obj = o_obj, q.tri = o_tri, q.t = o_t, q.u = o_u, q.v = o_v;
  q.any = d_out;
  FrameParams fp;
  std::memset(&fp, 0, sizeof(fp));
  for (int i = 0; i < 16; ++i) fp.cs[i] = (i % 5 == 0) ? 1.f : 0.f;
  int mode = any ? kModeAnyRays : kModeRays;
  if (n_slots > J3DG_MAX_OBJECTS) {
    float r = 0.f;
    for (uint64_t i = 0; i < 3 * n; ++i) {
      const float a = std::fabs(rays[8 * (i / 3) + i % 3]);
      if (std::isfinite(a)) r = std::max(r, a);  // a ray with a non-finite origin hits nothing
    }
    fp.top_nodes = ctx->d_top;
    fp.top_delta = ctx->top_delta(r);
    fp.top_count = ctx->top_count;
    mode |= kModeTop;
  }
  int rc = ctx->launch_trace_kernel(mode, fp, q, (uint32_t)n, n_slots, max_stack, width, ctx->d_queue);
  if (rc) return rc;
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  if (any) {
    HIP_CALL(hipMemcpy(hit, d_out, n, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  } else {
    if (obj) HIP_CALL(hipMemcpy(obj, o_obj, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (tri) HIP_CALL(hipMemcpy(tri, o_tri, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (t) HIP_CALL(hipMemcpy(t, o_t, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (u) HIP_CALL(hipMemcpy(u, o_u, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (v) HIP_CALL(hipMemcpy(v, o_v, n * 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  }
  uint64_t c[32];
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

// j3dg_trace_rays on device buffers with no allocation: the queue words,
// counters, object table and stack spill area live in this GPU's `scratch`
// array (noalloc_scratch, the address looked up at j3dg_create); the grid is
// cut to what the spill area holds.  Calls that overlap on one GPU share that
// array and race.
int j3dg_trace_rays_device(j3dg_ctx* h, const uint32_t* slots, uint32_t n_slots, const float* rays, uint64_t n,
                           uint32_t* obj, uint32_t* tri, float* t, float* u, float* v) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!ctx->noalloc_scratch) return ctx->fail(J3DG_ERR_LAUNCH, "no scratch address for this device");
  if (n_slots > J3DG_MAX_OBJECTS)
    return ctx->fail(J3DG_ERR_INVALID_ARG, "j3dg_trace_rays_device takes at most J3DG_MAX_OBJECTS mesh slots");
  if (n_slots && !slots) return ctx->fail(J3DG_ERR_INVALID_ARG, "null slot list");
  if (n > 0xFFFFFFFFull / 2) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many rays in one call");
  if (n == 0) return J3DG_OK;
  if (!rays || !obj || !tri || !t || !u || !v) return ctx->fail(J3DG_ERR_INVALID_ARG, "null device buffer");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  DevObject tab[J3DG_MAX_OBJECTS];
  std::memset(tab, 0, sizeof(tab));
  uint32_t max_stack = 1;
  int width = 4;
  if (n_slots) {
    int rc = ctx->fill_object_table(slots, n_slots, tab, max_stack, width);
    if (rc) return rc;
  }
  const uint32_t D = ctx->tuning.stack_depth, block = ctx->tuning.block_size;
  const uint32_t cap = max_stack > D ? max_stack - D + 1 : 1;
  const uint64_t spill_entries = (kNoAllocScratchBytes - kNoAllocSpill) / sizeof(uint2);
  const uint64_t blocks_fit = spill_entries / ((uint64_t)cap * block);
  if (blocks_fit == 0)
    return ctx->fail(J3DG_ERR_UNSUPPORTED, "the meshes' traversal-stack bound exceeds the scratch area (" +
                                               std::to_string(kNoAllocScratchBytes >> 20) + " MiB)");
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ctx->trace_grid((uint32_t)n, D, block), blocks_fit);
  unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
  uint32_t* queue = reinterpret_cast<uint32_t*>(base);
  uint64_t* counters = reinterpret_cast<uint64_t*>(base + kNoAllocCounters);
  DevObject* objs = reinterpret_cast<DevObject*>(base + kNoAllocObjects);
  uint2* spill = reinterpret_cast<uint2*>(base + kNoAllocSpill);
  // queue heads and counters start at zero: a fill kernel, no memset node
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ctx->stream, queue, counters, 32u);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  if (n_slots)
    HIP_CALL(hipMemcpyAsync(objs, tab, n_slots * sizeof(DevObject), hipMemcpyHostToDevice, ctx->stream),
             J3DG_ERR_LAUNCH);
  RayQuery q;
  std::memset(&q, 0, sizeof(q));
  q.rays = rays;
  q.obj = obj, q.tri = tri, q.t = t, q.u = u, q.v = v;
  FrameParams f;
  std::memset(&f, 0, sizeof(f));
  BatchParams bp;
  std::memset(&bp, 0, sizeof(bp));
  const DevObject* objs_c = objs;
  uint32_t nr = (uint32_t)n, no = n_slots, capv = cap;
  uint32_t batch = 64u * ctx->tuning.rays_per_item, refill = ctx->tuning.refill_min, coop = ctx->tuning.coop_tail;
  uint64_t* tl = nullptr;
  void* args[] = {&f, &q, &nr, &objs_c, &no, &spill, &capv, &queue, &counters, &batch, &refill, &coop, &tl, &bp};
  void* fn = j3d::trace_kernel(kModeRays, width, (int)D, false);
  HIP_CALL(hipLaunchKernel(fn, dim3(grid), dim3(block), args, D * block * sizeof(uint2), ctx->stream),
           J3DG_ERR_LAUNCH);
  uint64_t overflow = 0;
  HIP_CALL(hipMemcpyAsync(&overflow, counters + 6, sizeof(overflow), hipMemcpyDeviceToHost, ctx->stream),
           J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  if (overflow) return ctx->fail(J3DG_ERR_INTERNAL, "traversal stack overflow");
  return J3DG_OK;
}

int j3dg_calibration_copy(j3dg_ctx* h, uint64_t bytes, uint32_t repeats, float* ms) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (bytes < 16 || bytes % 16) return ctx->fail(J3DG_ERR_INVALID_ARG, "bytes must be a positive multiple of 16");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  float4 *a = nullptr, *b = nullptr;
  HIP_CALL(hipMalloc(&a, bytes), J3DG_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&b, bytes) != hipSuccess) {
    hipFree(a);
    return ctx->fail(J3DG_ERR_OUT_OF_MEMORY, "calibration buffer");
  }
  uint64_t n = bytes / 16;
  hipMemsetAsync(a, 1, bytes, ctx->stream);
  hipEvent_t e0 = ctx->ev[0], e1 = ctx->ev[1];
  hipEventRecord(e0, ctx->stream);
  for (uint32_t r = 0; r < std::max(repeats, 1u); ++r)
    hipLaunchKernelGGL(k_copy, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->n_cus * 16)), dim3(256),
                       0, ctx->stream, a, b, n);
  hipEventRecord(e1, ctx->stream);
  hipError_t err = hipStreamSynchronize(ctx->stream);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  if (ms) *ms = t / std::max(repeats, 1u);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) return ctx->fail(J3DG_ERR_LAUNCH, hipGetErrorString(err));
  return J3DG_OK;
}

int j3dg_debug_timeline(j3dg_ctx* h, uint64_t* out, uint32_t max_waves, uint32_t* n_waves) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!n_waves) return ctx->fail(J3DG_ERR_INVALID_ARG, "null count");
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  uint32_t n = std::min(max_waves, ctx->timeline_waves);
  if (n && out && ctx->d_timeline)
    HIP_CALL(hipMemcpy(out, ctx->d_timeline, (size_t)n * 8 * sizeof(uint64_t), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  *n_waves = n;
  return J3DG_OK;
}

// Many frames (cameras x meshes of uneven size) in one launch: per-frame
// parameters and object tables go to device tables, one persistent k_trace
// (kModeBatch) takes tiles of all frames from one queue, one k_shade_batch
// shades them; reset + trace + shade are captured in a HIP graph and replayed
// while the launch parameters are unchanged (the tables' contents may change).
// The batch launch reads the batch's object table (d_bobjs), and the
// cooperative tail hands a ray's frame (index, object table, object count) to
// the lanes that take it over together with its state.
static constexpr bool kBatchEnabled = true;

int j3dg_render_batch(j3dg_ctx* h, const j3dg_frame_desc* frames, const uint32_t* targets, uint32_t n) {
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  CHECK_CTX();
  if (!kBatchEnabled)
    return ctx->fail(J3DG_ERR_UNSUPPORTED, "j3dg_render_batch is disabled: its launch faulted the GPU and is under analysis");
  if (n == 0) return J3DG_OK;
  if (!frames || !targets) return ctx->fail(J3DG_ERR_INVALID_ARG, "null frame or target list");
  if (n > 65536) return ctx->fail(J3DG_ERR_INVALID_ARG, "more than 65536 frames in one batch");
  for (uint32_t i = 0; i < n; ++i) {  // surface colours come from the single-frame k_color pass
    if (frames[i].flags & (J3DG_TEXTURED | J3DG_VERTEXCOLORS))
      return ctx->fail(J3DG_ERR_UNSUPPORTED, "textured / vertex-coloured frames render one at a time");
    if (frames[i].object_list)
      return ctx->fail(J3DG_ERR_UNSUPPORTED, "frames with an object_list render one at a time");
  }
  if (!ctx->d_matcap) return ctx->fail(J3DG_ERR_INVALID_ARG, "no matcap set");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  std::vector<j3d::FrameParams> fps(n);
  std::vector<j3d::ShadeParams> sps(n);
  std::vector<uint32_t> first(n + 1, 0);
  std::vector<j3d::DevObject> objs;
  uint32_t max_stack = 1, max_w = 0, max_h = 0;
  int width = 0;
  std::vector<uint32_t> infer;  // learned frames whose proposals the network computes
  for (uint32_t i = 0; i < n; ++i) {
    const j3dg_frame_desc& fd = frames[i];
    if (targets[i] >= ctx->targets.size() || !ctx->targets[targets[i]].live)
      return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown target in batch");
    Target& tg = ctx->targets[targets[i]];
    if (fd.width != tg.w || fd.height != tg.h) return ctx->fail(J3DG_ERR_INVALID_ARG, "frame size differs from target size");
    if (fd.flags & J3DG_NO_SCENE) return ctx->fail(J3DG_ERR_UNSUPPORTED, "NO_SCENE frames render one at a time");
    if (fd.n_objects > J3DG_MAX_OBJECTS) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many objects");
    int rc = j3d::make_frame_params(fd, tg, fps[i]);
    if (rc) return ctx->fail(rc, "invalid frame in batch");
    fps[i].fused = 1;
    // the frame's objects, as a single frame's table holds them (networks included)
    const size_t base = objs.size();
    fps[i].obj_base = (uint32_t)base;
    objs.resize(base + fd.n_objects);
    if (fd.n_objects) std::memset(&objs[base], 0, fd.n_objects * sizeof(j3d::DevObject));
    uint32_t fstack = 0;
    int fwidth = 0;
    rc = ctx->fill_object_table(fd.objects, fd.n_objects, objs.data() + base, fstack, fwidth);
    if (rc) return rc;
    if (fd.n_objects) {
      if (width && width != fwidth) return ctx->fail(J3DG_ERR_UNSUPPORTED, "a batch's meshes must share BVH width");
      width = fwidth;
    }
    max_stack = std::max(max_stack, fstack);
    // learned frames: the bounded pass reads the target's proposal buffer,
    // filled by the network before the launch (or kept, J3DG_KEEP_PROPOSAL)
    if ((fd.flags & J3DG_LEARNED) && fd.n_objects) {
      bool any_net = false;
      for (uint32_t k = 0; k < fd.n_objects; ++k) any_net = any_net || ctx->meshes[fd.objects[k]].has_proposal;
      const bool keep = (fd.flags & J3DG_KEEP_PROPOSAL) != 0;
      if (keep || any_net) {
        rc = ensure_proposal(ctx, tg);
        if (rc) return rc;
        fps[i].proposal = tg.proposal;
        if (!keep) infer.push_back(i);
      }
    }
    first[i + 1] = first[i] + fps[i].n_tiles;
    j3d::make_shade_params(fd, tg, *ctx, sps[i]);
    max_w = std::max(max_w, tg.w), max_h = std::max(max_h, tg.h);
  }
  if (!width) width = 8;
  const uint32_t total_tiles = first[n];
  if ((uint64_t)total_tiles * 64u >= 0xFFFFFFFFull) return ctx->fail(J3DG_ERR_INVALID_ARG, "batch has too many pixels");
  // device tables (grown as needed; contents re-uploaded when they change)
  if (n > ctx->bframes_cap || objs.size() > ctx->bobjs_cap) {
    HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
    ctx->invalidate_graphs();
    if (n > ctx->bframes_cap) {
      if (ctx->d_bframes) hipFree(ctx->d_bframes);
      if (ctx->d_bfirst) hipFree(ctx->d_bfirst);
      if (ctx->d_bshade) hipFree(ctx->d_bshade);
      ctx->d_bframes = nullptr, ctx->d_bfirst = nullptr, ctx->d_bshade = nullptr;
      ctx->bframes_cap = 0;
      HIP_CALL(hipMalloc(&ctx->d_bframes, n * sizeof(j3d::FrameParams)), J3DG_ERR_OUT_OF_MEMORY);
      HIP_CALL(hipMalloc(&ctx->d_bfirst, (n + 1) * sizeof(uint32_t)), J3DG_ERR_OUT_OF_MEMORY);
      HIP_CALL(hipMalloc(&ctx->d_bshade, n * sizeof(j3d::ShadeParams)), J3DG_ERR_OUT_OF_MEMORY);
      ctx->bframes_cap = n;
    }
    if (objs.size() > ctx->bobjs_cap) {
      if (ctx->d_bobjs) hipFree(ctx->d_bobjs);
      ctx->d_bobjs = nullptr;
      ctx->bobjs_cap = 0;
      HIP_CALL(hipMalloc(&ctx->d_bobjs, std::max<size_t>(objs.size(), 1) * sizeof(j3d::DevObject)),
               J3DG_ERR_OUT_OF_MEMORY);
      ctx->bobjs_cap = (uint32_t)std::max<size_t>(objs.size(), 1);
    }
    ctx->batch_tables.clear();
  }
  {
    std::vector<uint8_t> tab;
    auto put = [&](const void* p, size_t b) {
      const uint8_t* q = reinterpret_cast<const uint8_t*>(p);
      tab.insert(tab.end(), q, q + b);
    };
    put(fps.data(), n * sizeof(j3d::FrameParams));
    put(first.data(), (n + 1) * sizeof(uint32_t));
    put(sps.data(), n * sizeof(j3d::ShadeParams));
    put(objs.data(), objs.size() * sizeof(j3d::DevObject));
    if (tab != ctx->batch_tables) {
      HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(ctx->d_bframes, fps.data(), n * sizeof(j3d::FrameParams), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(ctx->d_bfirst, first.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(ctx->d_bshade, sps.data(), n * sizeof(j3d::ShadeParams), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
      if (!objs.empty())
        HIP_CALL(hipMemcpy(ctx->d_bobjs, objs.data(), objs.size() * sizeof(j3d::DevObject), hipMemcpyHostToDevice),
                 J3DG_ERR_LAUNCH);
      ctx->batch_tables.swap(tab);
    }
  }
  // the trace launch's shape; the spill area must exist before capture
  const uint32_t D = ctx->tuning.stack_depth, block = ctx->tuning.block_size;
  const uint32_t n_rays = total_tiles * 64u;
  const uint32_t grid = ctx->trace_grid(n_rays, D, block);
  const uint32_t cap = max_stack > D ? max_stack - D + 1 : 1;
  int rc = ctx->ensure_spill((uint64_t)grid * block, cap);
  if (rc) return rc;
  j3d::BatchParams bp;
  bp.frames = ctx->d_bframes;
  bp.first_tile = ctx->d_bfirst;
  bp.n_frames = n;
  const dim3 sgrid((max_w + 63) / 64, (max_h + 3) / 4, n);
  struct Key {
    void* fn;
    uint32_t grid, block, D, n_rays, cap, n, sx, sy, rpi, refill, coop, count;
    const void *spill, *frames, *first, *shade, *objs;
  } key;
  std::memset(&key, 0, sizeof(key));
  key.fn = j3d::trace_kernel(j3d::kModeBatch, width, (int)D, ctx->counting);
  key.grid = grid, key.block = block, key.D = D, key.n_rays = n_rays, key.cap = cap, key.n = n;
  key.sx = sgrid.x, key.sy = sgrid.y, key.rpi = ctx->tuning.rays_per_item, key.refill = ctx->tuning.refill_min;
  key.coop = ctx->tuning.coop_tail, key.count = ctx->counting ? 1u : 0u;
  key.spill = ctx->spill, key.frames = ctx->d_bframes, key.first = ctx->d_bfirst, key.shade = ctx->d_bshade;
  key.objs = ctx->d_bobjs;
  std::vector<uint8_t> kb(reinterpret_cast<uint8_t*>(&key), reinterpret_cast<uint8_t*>(&key) + sizeof(key));
  if (!ctx->graph_exec || kb != ctx->graph_key) {
    ctx->invalidate_graphs();
    HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
    HIP_CALL(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), J3DG_ERR_LAUNCH);
    hipLaunchKernelGGL(j3d::k_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->d_queue, ctx->d_counters, 32u);
    j3d::FrameParams f0;
    std::memset(&f0, 0, sizeof(f0));
    j3d::RayQuery q;
    std::memset(&q, 0, sizeof(q));
    rc = ctx->launch_trace_kernel(j3d::kModeBatch, f0, q, n_rays, 0, max_stack, width, ctx->d_queue, &bp);
    if (rc == J3DG_OK) hipLaunchKernelGGL(j3d::k_shade_batch, sgrid, dim3(256), 0, ctx->stream, (const j3d::ShadeParams*)ctx->d_bshade);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    HIP_CALL(ce, J3DG_ERR_LAUNCH);
    ctx->graph = g;
    HIP_CALL(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0), J3DG_ERR_LAUNCH);
    ctx->graph_key.swap(kb);
  }
  // the learned frames' proposals, before the graph (they change with the cameras)
  for (size_t k = 0; k < infer.size(); ++k) {
    const uint32_t i = infer[k];
    rc = ctx->launch_mlp_frame(fps[i], fps[i].n_obj, ctx->d_bobjs + fps[i].obj_base, k == 0, k + 1 == infer.size());
    if (rc) return rc;
  }
  hipEvent_t* e = ctx->cur_ev();
  HIP_CALL(hipEventRecord(e[0], ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipGraphLaunch(ctx->graph_exec, ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipEventRecord(e[1], ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipEventRecord(e[6], ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipEventRecord(e[7], ctx->stream), J3DG_ERR_LAUNCH);
  ctx->graph_launches++;
  ctx->traced = true, ctx->shaded = false, ctx->mlp_ran = !infer.empty();
  ctx->last_ev = e;
  // trace time covers trace + shade; mlp time, every learned frame's network
  ctx->ring_flags[ctx->ring_n % Ctx::kRing] = infer.empty() ? 1u : 5u;
  ctx->ring_n++;
  return J3DG_OK;
}

}  // extern "C"
