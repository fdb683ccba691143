// mlp.hip -- the learned tmax proposal on gfx950.
//
// A 78-64-64-64-2 MLP over Fourier features of the object-space ray predicts
// the normalised distance to the closest hit and a hit logit.  The primary
// trace first traverses with tmax = prediction * (1 + margin): any hit inside
// that bound is the global closest hit, so the bounded pass is exact, and a
// ray that finds nothing falls back to the unbounded traversal.
//
// Network (the checker is oracle/mlp_oracle.cpp):
//   x   = ((o - c) / R, d / |d|)                     c, R: box centre, half-diagonal
//   f   = x, then sin/cos(2^k pi x_i), k = 0..5       (78 values, padded to 80)
//   h1..h3 = relu(W h + b) (64 each),  y = W4 h3 + b4 (y0 = |d| t / R, y1 = hit logit)
//
// Matrix layout.  v_mfma_f32_32x32x16_bf16 multiplies a 32x16 A by a 16x32 B:
// lane l (i = l % 32, half h = l / 32) holds A[i][k] and B[k][i] for the 8
// k of its half; the fp32 accumulator holds D[row][l % 32] for rows
// 8 (v / 4) + 4 h + v % 4, v = 0..15.  Rays are the 32 columns (N), neurons
// the rows (M), inputs the reduction (K).  Because the accumulator rows a lane
// holds are exactly the K values it must supply to the next layer's B operand
// if the K order is permuted (kperm below), a layer's output feeds the next
// layer straight from registers; the permutation is folded into the packed
// weight fragments instead.  Weight gradients (K = rays) need the transposed
// view and go through a per-wave LDS tile.
//
// All weights and activations enter the matrix cores as bf16 (round to
// nearest even), the 2-output layer included (a 32-row tile with two live
// rows); accumulators start at the bias; the loss, the output layer's
// backward, bias sums and Adam are fp32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "j3dg.h"
#include "j3dg_internal.h"
#include "traverse.h"

namespace j3d {
namespace mlp {

constexpr int kIn = 78;
constexpr int kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + 2 * kH, kParams = oB4 + 2;
static_assert(kParams == J3DG_MLP_PARAMS, "parameter count");
constexpr float kRegWeight = 10.f;  // weight of the distance term in the loss

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Weights as MFMA A fragments ([fragment][lane], 16 B per lane), biases and the
// output layer in accumulator order.  Built from the fp32 parameters by k_pack.
struct Packed {
  u32x4 f1[10][64];     // W1: fragment mt * 5 + kb
  u32x4 f2[8][64];      // W2: mt * 4 + kb
  u32x4 f3[8][64];      // W3
  u32x4 t2[8][64];      // W2^T (backward): rows = layer-2 inputs
  u32x4 t3[8][64];      // W3^T
  u32x4 f4[4][64];      // W4 as a 32-row tile (rows 2..31 zero): fragment kb
  float bc[3][2][32];   // b1..b3: [layer][half][mt * 16 + v]
  float w4c[2][2][32];  // bf16(W4): [output][half][mt * 16 + v]
  float b4[4];
};
constexpr int kFrags = 46;

// K position of slot s (0..7) of half h in K block kb.
__host__ __device__ constexpr int kperm(int kb, int h, int s) { return 16 * kb + 8 * (s >> 2) + 4 * h + (s & 3); }
// Row held by accumulator register v of half h in row tile mt.
__host__ __device__ constexpr int crow(int mt, int h, int v) { return 32 * mt + 8 * (v >> 2) + 4 * h + (v & 3); }

__device__ inline uint32_t pack2(float a, float b) {
  bf16x2 v;
  v.x = (__bf16)a;
  v.y = (__bf16)b;
  return __builtin_bit_cast(uint32_t, v);
}
__device__ inline float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
__device__ inline float bf_slot(const u32x4& q, int s) { return (s & 1) ? bf_hi(q[s >> 1]) : bf_lo(q[s >> 1]); }

__device__ inline f32x16 mfma(const u32x4& a, const u32x4& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

__device__ inline u32x4 ldfrag(const u32x4* p) { return *(const __attribute__((address_space(1))) u32x4*)p; }

// ---------------------------------------------------------------- packing ---

__global__ void k_pack(const float* __restrict__ p, Packed* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kFrags * 64) {
    const int fr = t / 64, l = t % 64, i = l & 31, h = l >> 5;
    float w[8];
    for (int s = 0; s < 8; ++s) {
      float v = 0.f;
      if (fr < 10) {
        const int mt = fr / 5, k = kperm(fr % 5, h, s);
        if (k < kIn) v = p[oW1 + (32 * mt + i) * kIn + k];
      } else if (fr < 26) {
        const int g = fr < 18 ? fr - 10 : fr - 18, off = fr < 18 ? oW2 : oW3;
        v = p[off + (32 * (g / 4) + i) * kH + kperm(g % 4, h, s)];
      } else if (fr < 42) {
        const int g = fr < 34 ? fr - 26 : fr - 34, off = fr < 34 ? oW2 : oW3;
        v = p[off + kperm(g % 4, h, s) * kH + 32 * (g / 4) + i];
      } else if (i < 2) {
        v = p[oW4 + i * kH + kperm(fr - 42, h, s)];
      }
      w[s] = v;
    }
    u32x4 r = {pack2(w[0], w[1]), pack2(w[2], w[3]), pack2(w[4], w[5]), pack2(w[6], w[7])};
    reinterpret_cast<u32x4*>(out)[t] = r;
    return;
  }
  int u = t - kFrags * 64;
  if (u < 192) {
    const int layer = u / 64, h = (u / 32) & 1, c = u & 31;
    const int off = layer == 0 ? oB1 : layer == 1 ? oB2 : oB3;
    out->bc[layer][h][c] = p[off + crow(c / 16, h, c % 16)];
    return;
  }
  u -= 192;
  if (u < 128) {
    const int o = u / 64, h = (u / 32) & 1, c = u & 31;
    out->w4c[o][h][c] = (float)(__bf16)p[oW4 + o * kH + crow(c / 16, h, c % 16)];
    return;
  }
  u -= 128;
  if (u < 4) out->b4[u] = u < 2 ? p[oB4 + u] : 0.f;
}
constexpr int kPackThreads = kFrags * 64 + 192 + 128 + 4;

// -------------------------------------------------------------- features ---

// Network input of one object-space ray, as the line it lies on: the line's
// point closest to the box centre c (offset from c, over R) and the unit
// direction u.  Returns |d|; a = (c - o) . u is the distance from o to that
// point, so a prediction y0 maps back to t = (y0 R + a) / |d|.
__host__ __device__ inline float ray_input(const float* nrm, float ox, float oy, float oz, float dx, float dy,
                                           float dz, float x[6], float& a) {
  const float dl = sqrtf(dx * dx + dy * dy + dz * dz);
  const float rl = 1.f / dl, rR = 1.f / nrm[3];
  const float ux = dx * rl, uy = dy * rl, uz = dz * rl;
  a = (nrm[0] - ox) * ux + (nrm[1] - oy) * uy + (nrm[2] - oz) * uz;
  x[0] = ((ox + a * ux) - nrm[0]) * rR;
  x[1] = ((oy + a * uy) - nrm[1]) * rR;
  x[2] = ((oz + a * uz) - nrm[2]) * rR;
  x[3] = ux;
  x[4] = uy;
  x[5] = uz;
  return dl;
}

// Feature idx (even) and idx + 1: kind 0 = the raw inputs, 1 = sin/cos pair, 2 = padding.
__device__ constexpr int feat_kind(int idx) { return idx < 6 ? 0 : idx < kIn ? 1 : 2; }
// sin/cos pair idx: angle 2^k pi x_i in revolutions (x_i * 2^(k-1)), exact scaling.
__device__ inline float feat_rev(const float x[6], int idx) {
  const int m = (idx - 6) >> 1;
  return x[m / 6] * (float)(1 << (m % 6)) * 0.5f;
}
__device__ inline void feat_static(const float x[6], int idx, float& a, float& b) {
  const int kind = feat_kind(idx);
  if (kind == 0) {
    a = x[idx];
    b = x[idx + 1];
  } else if (kind == 1) {
    const float r = feat_rev(x, idx);
    a = __builtin_amdgcn_sinf(r);
    b = __builtin_amdgcn_cosf(r);
  } else {
    a = b = 0.f;
  }
}

// The 40 features lane half h supplies, as the 5 B fragments of layer 1.
__device__ inline void features(const float x[6], uint32_t h, u32x4 fr[5]) {
#pragma unroll
  for (int kb = 0; kb < 5; ++kb) {
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) {
      const int i0 = kperm(kb, 0, 2 * sp), i1 = kperm(kb, 1, 2 * sp);
      float a, b;
      if (feat_kind(i0) == 1 && feat_kind(i1) == 1) {
        const float r = h ? feat_rev(x, i1) : feat_rev(x, i0);
        a = __builtin_amdgcn_sinf(r);
        b = __builtin_amdgcn_cosf(r);
      } else {
        float a0, b0, a1, b1;
        feat_static(x, i0, a0, b0);
        feat_static(x, i1, a1, b1);
        a = h ? a1 : a0;
        b = h ? b1 : b0;
      }
      fr[kb][sp] = pack2(a, b);
    }
  }
}

// ---------------------------------------------------------------- layers ---

// Pre-activations (2 row tiles, bias included) -> relu -> the next layer's 4
// B fragments; mask bit (mt * 16 + v) = pre-activation > 0.
__device__ inline void activate(const f32x16 c[2], u32x4 out[4], uint32_t& mask) {
  float v[32];
  mask = 0;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float z = c[mt][r];
      const bool pos = z > 0.f;
      mask |= (pos ? 1u : 0u) << (mt * 16 + r);
      v[mt * 16 + r] = pos ? z : 0.f;
    }
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int w = 0; w < 4; ++w) out[kb][w] = pack2(v[8 * kb + 2 * w], v[8 * kb + 2 * w + 1]);
}

// fp32 values in accumulator order -> 4 B fragments (bf16).
__device__ inline void to_frags(const float v[32], u32x4 out[4]) {
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int w = 0; w < 4; ++w) out[kb][w] = pack2(v[8 * kb + 2 * w], v[8 * kb + 2 * w + 1]);
}

// c = W in + b: the accumulators start at the bias (accumulator order).
__device__ inline f32x16 bias_tile(const float* bias, int mt) {
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = bias[mt * 16 + r];
  return acc;
}

template <int KB>
__device__ inline void dense(const u32x4* a /* [2 * KB] */, const float* bias, const u32x4* in, f32x16 c[2]) {
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16 acc = bias_tile(bias, mt);
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) acc = mfma(a[mt * KB + kb], in[kb], acc);
    c[mt] = acc;
  }
}

// Output layer: one more row tile whose rows 0 and 1 are W4 (the rest zero);
// y0, y1 land in registers 0 and 1 of the lane half h = 0 and are broadcast
// to the other half.
__device__ inline void output_layer(const Packed* pk, const u32x4* a4, uint32_t lane, const u32x4 h3[4], float& y0,
                                    float& y1) {
  const bool h0 = lane < 32;
  f32x16 acc = {};
  acc[0] = h0 ? pk->b4[0] : 0.f;
  acc[1] = h0 ? pk->b4[1] : 0.f;
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = mfma(a4[kb], h3[kb], acc);
  y0 = __shfl(acc[0], (int)(lane & 31u), 64);
  y1 = __shfl(acc[1], (int)(lane & 31u), 64);
}

// Forward pass of one chunk of 32 rays (both halves of the wave hold the same ray).
struct Weights {
  u32x4 a1[10], a2[8], a3[8], a4[4];
};
__device__ inline void load_weights(const Packed* pk, uint32_t lane, Weights& w) {
#pragma unroll
  for (int i = 0; i < 10; ++i) w.a1[i] = ldfrag(&pk->f1[i][lane]);
#pragma unroll
  for (int i = 0; i < 8; ++i) w.a2[i] = ldfrag(&pk->f2[i][lane]);
#pragma unroll
  for (int i = 0; i < 8; ++i) w.a3[i] = ldfrag(&pk->f3[i][lane]);
#pragma unroll
  for (int i = 0; i < 4; ++i) w.a4[i] = ldfrag(&pk->f4[i][lane]);
}

__device__ inline void forward(const Packed* pk, const Weights& w, uint32_t lane, const float x[6], float& y0,
                               float& y1) {
  const uint32_t h = lane >> 5;
  u32x4 fin[5], h1[4], h2[4], h3[4];
  uint32_t m;
  f32x16 c[2];
  features(x, h, fin);
  dense<5>(w.a1, pk->bc[0][h], fin, c);
  activate(c, h1, m);
  dense<4>(w.a2, pk->bc[1][h], h1, c);
  activate(c, h2, m);
  dense<4>(w.a3, pk->bc[2][h], h2, c);
  activate(c, h3, m);
  output_layer(pk, w.a4, lane, h3, y0, y1);
}

// -------------------------------------------------------------- inference ---

constexpr uint32_t kInferQueue = 128;  // ring of queued pixels per wave: < 32 left + 64 new fit

// RAW: network outputs (y0, y1) of object-space rays (n x 8 floats).
// Frame: per pixel of the rectangle, the bounded-pass tmax and hit logit over
// the frame's objects that carry a network: (min of t_pred * (1 + margin),
// max logit) over the objects predicting a hit; logit -1 if none does.
template <bool RAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_infer(FrameParams f, const DevObject* __restrict__ objs, uint32_t n_obj,
                                               const Packed* raw_pk, float4 raw_norm, const float* __restrict__ rays,
                                               uint32_t n_rays, float2* __restrict__ out) {
  const uint32_t lane0 = threadIdx.x & 63u;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t n_items = RAW ? n_rays : f.n_tiles * 64u;
  const uint32_t n_chunks = (n_items + 31u) / 32u;
  const float t_near = f.diagonal / 100.f;
  // frames: per wave, a ring of pixel indices waiting for the network
  __shared__ uint32_t q_lds[RAW ? 1 : 4][RAW ? 1 : kInferQueue];
  bool first = true;
  for (uint32_t o = 0; o < (RAW ? 1u : n_obj); ++o) {
    const Packed* pk = RAW ? raw_pk : reinterpret_cast<const Packed*>(objs[o].mlp);
    if (!pk) continue;
    float nrm[4];
    if (RAW) {
      nrm[0] = raw_norm.x, nrm[1] = raw_norm.y, nrm[2] = raw_norm.z, nrm[3] = raw_norm.w;
    } else {
      for (int k = 0; k < 4; ++k) nrm[k] = objs[o].mlp_norm[k];
    }
    Weights w;
    load_weights(pk, lane0, w);
    uint32_t lane = lane0;
    if constexpr (RAW) {
      for (uint32_t c = wave; c < n_chunks; c += n_waves) {
        // opaque per-trip lane index: keeps lane-dependent constants inside the loop
        asm volatile("" : "+v"(lane));
        const uint32_t j = lane & 31u;
        const uint32_t id = c * 32u + j;
        const bool valid = id < n_items;
        f4 ro = {0.f, 0.f, 0.f, 1.f}, rd = {0.f, 0.f, 1.f, 0.f};
        if (valid) {
          const float* rr = rays + 8 * (size_t)id;
          ro = f4{rr[0], rr[1], rr[2], 1.f};
          rd = f4{rr[3], rr[4], rr[5], 0.f};
        }
        float x[6], a;
        ray_input(nrm, ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, x, a);
        float y0, y1;
        forward(pk, w, lane, x, y0, y1);
        if ((lane >> 5) == 0 && valid) out[id] = make_float2(y0, y1);
      }
    } else {
      // A ray that enters none of the object's root children boxes (the
      // traversal's own first test) cannot hit the object: it gets no
      // proposal from it (a miss proposal only sends a pixel to the plain
      // pass, which is exact for any proposal).  The other rays queue up in
      // LDS, 32 per network pass.
      uint32_t* qw = q_lds[threadIdx.x >> 6];
      uint32_t head = 0, tail = 0;  // wave-uniform
      const float* ic = objs[o].inv_cs;
      for (uint32_t b = wave;; b += n_waves) {
        const bool more = b < f.n_tiles;  // a tile's 64 pixels per pass of this loop
        if (more) {
          uint32_t px, py;
          bool need = false;
          if (tile_pixel(f, b, lane0, px, py)) {
            f4 wo, wd;
            primary_ray(f, px, py, wo, wd);
            const Ray r = make_object_ray(ic, wo, wd, t_near, FLT_MAX);
            need = objs[o].width == 8 ? enters_node<8>(objs[o].nodes, r, FLT_MAX)
                                      : enters_node<4>(objs[o].nodes, r, FLT_MAX);
            if (!need && first) out[(size_t)py * f.w + px] = make_float2(0.f, -1.f);
          }
          const uint64_t m = __ballot(need);
          if (need) qw[(tail + rank_in_mask(m)) & (kInferQueue - 1u)] = py * f.w + px;
          tail += (uint32_t)__popcll(m);
        }
        // the queue's writes reach LDS before any lane reads them back
        __builtin_amdgcn_wave_barrier();
        while (tail - head >= 32u || (!more && tail != head)) {
          asm volatile("" : "+v"(lane));
          const uint32_t j = lane & 31u;
          const bool valid = head + j < tail;
          const uint32_t pix = valid ? qw[(head + j) & (kInferQueue - 1u)] : 0u;
          f4 ro = {0.f, 0.f, 0.f, 1.f}, rd = {0.f, 0.f, 1.f, 0.f};
          if (valid) {
            f4 wo, wd;
            primary_ray(f, pix % f.w, pix / f.w, wo, wd);
            ro = mat_vec(ic, wo);
            rd = mat_vec(ic, wd);
          }
          float x[6], a;
          const float dl = ray_input(nrm, ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, x, a);
          float y0, y1;
          forward(pk, w, lane, x, y0, y1);
          if ((lane >> 5) == 0 && valid) {
            const float tb = ((y0 * nrm[3] + a) / dl) * (1.f + f.margin);
            const bool hit = y1 > 0.f && tb >= t_near && tb <= FLT_MAX;
            float2 cur = first ? make_float2(0.f, -1.f) : out[pix];
            if (hit) cur = cur.y > 0.f ? make_float2(fminf(cur.x, tb), fmaxf(cur.y, y1)) : make_float2(tb, y1);
            out[pix] = cur;
          }
          head += min(32u, tail - head);
          // every lane has read its entry before later passes overwrite the ring
          __builtin_amdgcn_wave_barrier();
        }
        if (!more) break;
      }
    }
    first = false;
  }
}

// --------------------------------------------------------------- training ---
//
// A block of 4 waves takes 4 chunks of 32 rays per trip.  Each wave runs the
// forward pass and the backward chain (deltas through W^T on the matrix
// cores) for its own chunk and leaves, in LDS as bf16 [neuron][ray] tiles, each
// layer's input and delta.  The 14 weight-gradient tiles (32 x 32, K = rays)
// are then split over the 4 waves, each accumulating its tiles over the
// block's 128 rays: that keeps the accumulators per wave at 4 tiles.

constexpr int kRS = 40;  // row stride of the LDS tiles in bf16 (80 B: 16-byte aligned, 2-way bank pattern)
struct ChunkTiles {
  uint16_t d[3][64 * kRS];  // deltas of layers 1..3
  uint16_t in0[80 * kRS];   // layer-1 input (features)
  uint16_t in1[64 * kRS];   // h1
  uint16_t in2[64 * kRS];   // h2
  uint16_t in3[64 * kRS];   // h3 (output layer, VALU)
  float dy[2][32];
};
constexpr int kGradTiles = 14;  // gW1: 2 x 3, gW2: 2 x 2, gW3: 2 x 2

__device__ inline uint16_t bf_bits(float x) { return (uint16_t)(pack2(x, 0.f) & 0xFFFFu); }

// Transposed bf16 store of a half's 32 accumulator-order values / B fragments.
__device__ inline void store_t(uint16_t* tile, uint32_t h, uint32_t j, const float v[32]) {
#pragma unroll
  for (int idx = 0; idx < 32; ++idx) tile[crow(idx / 16, h, idx % 16) * kRS + j] = bf_bits(v[idx]);
}
template <int NKB>
__device__ inline void store_t(uint16_t* tile, uint32_t h, uint32_t j, const u32x4* fr) {
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t word = fr[kb][w];
      tile[kperm(kb, h, 2 * w) * kRS + j] = (uint16_t)(word & 0xFFFFu);
      tile[kperm(kb, h, 2 * w + 1) * kRS + j] = (uint16_t)(word >> 16);
    }
}
// Row `row`, rays 16 kk + 8 h .. + 7 of a tile as a fragment.
__device__ inline u32x4 load_t(const uint16_t* tile, int row, int kk, uint32_t h) {
  return *reinterpret_cast<const u32x4*>(tile + row * kRS + 16 * kk + 8 * h);
}
__device__ inline float frag_sum(const u32x4& q) {
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s = s + bf_slot(q, k);
  return s;
}

// dH = W^T bf16(d) (accumulator order in and out); out = mask ? dH : 0.
__device__ inline void backprop(const u32x4 (*t)[64], uint32_t lane, const float d[32], uint32_t mask, float out[32]) {
  u32x4 dfr[4];
  to_frags(d, dfr);
  __builtin_amdgcn_sched_barrier(0);
  f32x16 c[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16 acc = {};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(ldfrag(&t[mt * 4 + kb][lane]), dfr[kb], acc);
    c[mt] = acc;
  }
#pragma unroll
  for (int idx = 0; idx < 32; ++idx) out[idx] = ((mask >> idx) & 1u) ? c[idx / 16][idx % 16] : 0.f;
}

// Forward with per-use weight loads (the accumulators need the registers).
template <int KB>
__device__ inline void dense_g(const u32x4 (*a)[64], const float* bias, uint32_t lane, const u32x4* in, f32x16 c[2]) {
  __builtin_amdgcn_sched_barrier(0);  // keep each layer's weight loads next to their use
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f32x16 acc = bias_tile(bias, mt);
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) acc = mfma(ldfrag(&a[mt * KB + kb][lane]), in[kb], acc);
    c[mt] = acc;
  }
}

// One gradient pass over n rays: block b takes `iters` trips of 4 chunks and
// writes its summed gradient (parameter order) to partial[b] and its loss sum
// to loss_out[b].  Gradients are of sum(loss) * inv_n.
__global__ __launch_bounds__(256, 1) void k_grad(const Packed* __restrict__ pk_arg, float4 norm4,
                                                 const float* __restrict__ rays, const float* __restrict__ labels,
                                                 uint32_t n, uint32_t iters, float inv_n, float* __restrict__ partial,
                                                 float* __restrict__ loss_out) {
  __shared__ ChunkTiles tiles[4];
  const uint32_t lane0 = threadIdx.x & 63u, wib0 = threadIdx.x >> 6;

  f32x16 acc[4];
  float gbs[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x16{};
  float gw4[2] = {0.f, 0.f}, gb4 = 0.f, loss = 0.f;

  uint32_t lane = lane0, wib = wib0;
  for (uint32_t it = 0; it < iters; ++it) {
    // Opaque per-trip copies: without them the compiler hoists hundreds of
    // lane-dependent constants and addresses out of the loop and spills them.
    const Packed* __restrict__ pk = pk_arg;
    asm volatile("" : "+v"(lane), "+v"(wib), "+s"(pk));
    const uint32_t j = lane & 31u, h = lane >> 5;
    ChunkTiles& T = tiles[wib];
    const float nrm[4] = {norm4.x, norm4.y, norm4.z, norm4.w};
    const uint32_t r = ((blockIdx.x * iters + it) * 4u + wib) * 32u + j;
    const bool valid = r < n;
    float ro[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 1.f};
    float s_lab = 0.f, hit = 0.f;
    if (valid) {
      for (int k = 0; k < 6; ++k) ro[k] = rays[8 * (size_t)r + k];
      s_lab = labels[2 * (size_t)r];
      hit = labels[2 * (size_t)r + 1];
    }
    float x[6], a_unused;
    ray_input(nrm, ro[0], ro[1], ro[2], ro[3], ro[4], ro[5], x, a_unused);
    u32x4 fin[5], h1[4], h2[4], h3[4];
    uint32_t m1, m2, m3;
    f32x16 c[2];
    features(x, h, fin);
    dense_g<5>(pk->f1, pk->bc[0][h], lane, fin, c);
    activate(c, h1, m1);
    dense_g<4>(pk->f2, pk->bc[1][h], lane, h1, c);
    activate(c, h2, m2);
    dense_g<4>(pk->f3, pk->bc[2][h], lane, h2, c);
    activate(c, h3, m3);
    float y0, y1;
    {
      __builtin_amdgcn_sched_barrier(0);
      u32x4 a4[4];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) a4[kb] = ldfrag(&pk->f4[kb][lane]);
      output_layer(pk, a4, lane, h3, y0, y1);
    }
    // loss: hit * 10 (y0 - s)^2 + softplus(y1) - hit * y1
    const float e = y0 - s_lab;
    const float dy0 = valid ? 2.f * kRegWeight * hit * e * inv_n : 0.f;
    const float dy1 = valid ? (1.f / (1.f + expf(-y1)) - hit) * inv_n : 0.f;
    if (valid && h == 0) loss += hit * kRegWeight * e * e + (fmaxf(y1, 0.f) + log1pf(expf(-fabsf(y1)))) - hit * y1;
    // deltas: d3 = relu'(h3) (bf16(W4)^T dy), then through W3^T and W2^T
    float d3[32], d2[32], d1[32];
#pragma unroll
    for (int idx = 0; idx < 32; ++idx) {
      const float dh = pk->w4c[0][h][idx] * dy0 + pk->w4c[1][h][idx] * dy1;
      d3[idx] = ((m3 >> idx) & 1u) ? dh : 0.f;
    }
    backprop(pk->t3, lane, d3, m2, d2);
    backprop(pk->t2, lane, d2, m1, d1);

    __syncthreads();  // the previous trip's readers are done with the tiles
    store_t<5>(T.in0, h, j, fin);
    store_t<4>(T.in1, h, j, h1);
    store_t<4>(T.in2, h, j, h2);
    store_t<4>(T.in3, h, j, h3);
    store_t(T.d[0], h, j, d1);
    store_t(T.d[1], h, j, d2);
    store_t(T.d[2], h, j, d3);
    if (h == 0) {
      T.dy[0][j] = dy0;
      T.dy[1][j] = dy1;
    }
    __syncthreads();

    // output layer (own chunk, lane = neuron): gW4[o][n] += sum_r dy_o(r) h3[n](r)
    {
      float a0 = 0.f, a1 = 0.f;
      for (int rr = 0; rr < 32; ++rr) {
        const float hv = __uint_as_float((uint32_t)T.in3[lane * kRS + rr] << 16);
        a0 = a0 + T.dy[0][rr] * hv;
        a1 = a1 + T.dy[1][rr] * hv;
      }
      gw4[0] += a0;
      gw4[1] += a1;
      if (lane < 2) {
        float sb = 0.f;
        for (int rr = 0; rr < 32; ++rr) sb = sb + T.dy[lane][rr];
        gb4 += sb;
      }
    }
    // this wave's weight-gradient tiles over the block's 4 chunks
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int tid = (int)wib + 4 * t;
      if (tid < kGradTiles) {
        const int layer = tid < 6 ? 0 : tid < 10 ? 1 : 2;
        const int local = tid < 6 ? tid : tid < 10 ? tid - 6 : tid - 10;
        const int nts = layer == 0 ? 3 : 2;
        const int mt = local / nts, nt = local % nts;
        for (int cw = 0; cw < 4; ++cw) {
          const ChunkTiles& C = tiles[cw];
          const uint16_t* in = layer == 0 ? C.in0 : layer == 1 ? C.in1 : C.in2;
#pragma unroll
          for (int kk = 0; kk < 2; ++kk) {
            const u32x4 a = load_t(C.d[layer], 32 * mt + (int)j, kk, h);
            const int brow = 32 * nt + (int)j;
            const u32x4 b = brow < (layer == 0 ? 80 : 64) ? load_t(in, brow, kk, h) : u32x4{0u, 0u, 0u, 0u};
            acc[t] = mfma(a, b, acc[t]);
            if (nt == 0) gbs[t] += frag_sum(a);
          }
        }
      }
    }
  }

  // write this block's gradient
  const uint32_t j = lane & 31u, h = lane >> 5;
  float* P = partial + (size_t)blockIdx.x * kParams;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int tid = (int)wib + 4 * t;
    if (tid < kGradTiles) {
      const int layer = tid < 6 ? 0 : tid < 10 ? 1 : 2;
      const int local = tid < 6 ? tid : tid < 10 ? tid - 6 : tid - 10;
      const int nts = layer == 0 ? 3 : 2;
      const int mt = local / nts, nt = local % nts;
      const int fan = layer == 0 ? kIn : kH;
      const int woff = layer == 0 ? oW1 : layer == 1 ? oW2 : oW3;
      const int boff = layer == 0 ? oB1 : layer == 1 ? oB2 : oB3;
      const int in = 32 * nt + (int)j;
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if (in < fan) P[woff + crow(mt, (int)h, v) * fan + in] = acc[t][v];
      const float sb = gbs[t] + __shfl_xor(gbs[t], 32, 64);
      if (nt == 0 && h == 0) P[boff + 32 * mt + (int)j] = sb;
    }
  }
  // output layer and loss: the 4 waves' sums in a fixed order
  __syncthreads();
  float* red = reinterpret_cast<float*>(&tiles[0]);  // [wave][2 * 64 + 2 + 1]
  float* mine = red + wib * 131;
  mine[lane] = gw4[0];
  mine[64 + lane] = gw4[1];
  if (lane < 2) mine[128 + lane] = gb4;
  for (int off = 32; off >= 1; off >>= 1) loss += __shfl_xor(loss, off, 64);
  if (lane == 0) mine[130] = loss;
  __syncthreads();
  if (wib == 0) {
    for (int k = (int)lane; k < 131; k += 64) {
      const float s = ((red[k] + red[131 + k]) + red[262 + k]) + red[393 + k];
      if (k < 128)
        P[oW4 + k] = s;
      else if (k < 130)
        P[oB4 + (k - 128)] = s;
      else
        loss_out[blockIdx.x] = s;
    }
  }
}

// Sum of the per-block partial gradients (fixed order), optionally an Adam step.
__global__ void k_update(const float* __restrict__ partial, uint32_t waves, const float* __restrict__ loss_part,
                         float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                         float* __restrict__ grad_out, float* __restrict__ loss_out, float loss_scale, float lr,
                         float c1, float c2, int update) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kParams) {
    float g = 0.f;
    for (uint32_t w = 0; w < waves; ++w) g += partial[(size_t)w * kParams + p];
    if (grad_out) grad_out[p] = g;
    if (update) {
      // beta1 = 0.9, beta2 = 0.999; 1 - beta as its own literal (1.f - 0.999f is 1.3e-5 off 0.001)
      const float b1 = 0.9f, b2 = 0.999f, ob1 = 0.1f, ob2 = 0.001f, eps = 1e-8f;
      const float mm = b1 * m[p] + ob1 * g;
      const float vv = b2 * v[p] + ob2 * g * g;
      m[p] = mm;
      v[p] = vv;
      params[p] = params[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
  }
  if (p == 0 && loss_out) {
    float s = 0.f;
    for (uint32_t w = 0; w < waves; ++w) s += loss_part[w];
    *loss_out = s * loss_scale;
  }
}

// ------------------------------------------------------------ training data ---

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline float urand(uint64_t seed, uint64_t i, uint32_t k) {
  const uint64_t z = mix64(seed ^ mix64(i * 16ull + k));
  return (float)(z >> 40) * (1.f / 16777216.f);
}
__device__ inline float3 normalize3(float3 a) {
  const float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  return make_float3(a.x / l, a.y / l, a.z / l);
}
__device__ inline float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Rays of random cameras around the box (object space, unit directions):
// camera c = i / per_cam on a sphere of radius R * U[dmin, dmax] around the
// centre, aimed at a uniform point of the box (half extents `half`) with a
// random roll; the ray through a uniform point of a vertical fov / aspect
// frustum.
__global__ void k_train_rays(uint64_t seed, uint32_t n, float4 norm4, float3 half, uint32_t per_cam, float dmin,
                             float dmax, float tan_half, float aspect, float* __restrict__ rays) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t cam = i / per_cam;
  const float R = norm4.w;
  const float z = 2.f * urand(seed, cam, 0) - 1.f, phi = 6.2831853f * urand(seed, cam, 1);
  const float rr = sqrtf(fmaxf(0.f, 1.f - z * z));
  const float dist = R * (dmin + (dmax - dmin) * urand(seed, cam, 2));
  const float3 eye = make_float3(norm4.x + rr * cosf(phi) * dist, norm4.y + rr * sinf(phi) * dist, norm4.z + z * dist);
  const float3 tgt = make_float3(norm4.x + half.x * (2.f * urand(seed, cam, 3) - 1.f),
                                 norm4.y + half.y * (2.f * urand(seed, cam, 4) - 1.f),
                                 norm4.z + half.z * (2.f * urand(seed, cam, 5) - 1.f));
  const float3 fwd = normalize3(make_float3(tgt.x - eye.x, tgt.y - eye.y, tgt.z - eye.z));
  float3 helper = make_float3(2.f * urand(seed, cam, 6) - 1.f, 2.f * urand(seed, cam, 7) - 1.f,
                              2.f * urand(seed, cam, 8) - 1.f);
  float3 right = cross3(fwd, helper);
  if (right.x * right.x + right.y * right.y + right.z * right.z < 1e-6f) right = cross3(fwd, make_float3(0.f, 0.f, 1.f));
  if (right.x * right.x + right.y * right.y + right.z * right.z < 1e-6f) right = cross3(fwd, make_float3(1.f, 0.f, 0.f));
  right = normalize3(right);
  const float3 up = cross3(right, fwd);
  const float a = (2.f * urand(seed, i, 9) - 1.f) * tan_half * aspect, b = (2.f * urand(seed, i, 10) - 1.f) * tan_half;
  const float3 d = normalize3(make_float3(fwd.x + a * right.x + b * up.x, fwd.y + a * right.y + b * up.y,
                                          fwd.z + a * right.z + b * up.z));
  float* o = rays + 8 * (size_t)i;
  o[0] = eye.x, o[1] = eye.y, o[2] = eye.z;
  o[3] = d.x, o[4] = d.y, o[5] = d.z;
  o[6] = 0.f;
  o[7] = FLT_MAX;
}

// Labels from the plain traversal: ((t |d| - a) / R, hit).
__global__ void k_train_labels(const float* __restrict__ rays, const uint32_t* __restrict__ obj,
                               const float* __restrict__ t, uint32_t n, float4 norm4, float* __restrict__ labels,
                               uint32_t* __restrict__ n_hits) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (i < n) {
    hit = obj[i] != kMiss;
    const float* r = rays + 8 * (size_t)i;
    const float nrm[4] = {norm4.x, norm4.y, norm4.z, norm4.w};
    float x[6], a;
    const float dl = ray_input(nrm, r[0], r[1], r[2], r[3], r[4], r[5], x, a);
    labels[2 * (size_t)i] = hit ? (t[i] * dl - a) / norm4.w : 0.f;
    labels[2 * (size_t)i + 1] = hit ? 1.f : 0.f;
  }
  // one atomic per block (per-wave atomics on one address serialise)
  __shared__ uint32_t wave_hits[4];  // launched with 256 threads
  const uint64_t b = __ballot(hit);
  if ((threadIdx.x & 63u) == 0) wave_hits[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t c = wave_hits[0] + wave_hits[1] + wave_hits[2] + wave_hits[3];
    if (c) atomicAdd(n_hits, c);
  }
}

}  // namespace mlp

// ------------------------------------------------------------------ host ---

using namespace mlp;

namespace {

uint32_t infer_grid(int n_cus, uint32_t n_items) {
  const uint32_t chunks = (n_items + 31u) / 32u;
  const uint32_t waves = std::max(1u, std::min(chunks, (uint32_t)n_cus * 8u));
  return (waves + 3u) / 4u;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// He-uniform weights, zero biases; the distance output starts at `s0`.
void init_params(uint64_t seed, float s0, std::vector<float>& p) {
  p.assign(kParams, 0.f);
  uint64_t st = seed ^ 0x5EEDull;
  auto fill = [&](int off, int fan_out, int fan_in, float scale) {
    const float lim = scale * std::sqrt(6.f / (float)fan_in);
    for (int i = 0; i < fan_out * fan_in; ++i) {
      const float u = (float)(splitmix(st) >> 40) * (1.f / 16777216.f);
      p[off + i] = (2.f * u - 1.f) * lim;
    }
  };
  fill(oW1, kH, kIn, 1.f);
  fill(oW2, kH, kH, 1.f);
  fill(oW3, kH, kH, 1.f);
  fill(oW4, 2, kH, 0.1f);
  p[oB4] = s0;
}

}  // namespace

int Ctx::launch_mlp_frame(const FrameParams& fp, uint32_t n_obj, const DevObject* objs, bool first, bool last) {
  Ctx* ctx = this;
  hipEvent_t* e = cur_ev();
  if (first) HIP_TRY(hipEventRecord(e[4], stream), J3DG_ERR_LAUNCH);
  const uint32_t grid = infer_grid(n_cus, fp.n_tiles * 64u);
  hipLaunchKernelGGL(k_infer<false>, dim3(grid), dim3(256), 0, stream, fp, objs ? objs : (const DevObject*)d_objs, n_obj,
                     (const Packed*)nullptr, make_float4(0.f, 0.f, 0.f, 1.f), (const float*)nullptr, 0u,
                     const_cast<float2*>(fp.proposal));
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  if (last) HIP_TRY(hipEventRecord(e[5], stream), J3DG_ERR_LAUNCH);
  mlp_ran = true;
  return J3DG_OK;
}

namespace {

// Gradient pass (+ optional Adam step) over n rays already on the device;
// `partial` holds blocks_cap * J3DG_MLP_PARAMS floats, loss_part blocks_cap.
int grad_pass(Ctx* ctx, const Packed* pk, const float norm[4], const float* rays, const float* labels, uint32_t n,
              float* partial, float* loss_part, uint32_t blocks_cap, float* params, float* m, float* v, float* grad_out,
              float* loss_out, float lr, float c1, float c2, bool update) {
  const uint32_t chunks = (n + 31u) / 32u;
  const uint32_t groups = (chunks + 3u) / 4u;  // one block trip = 4 chunks
  const uint32_t blocks = std::max(1u, std::min(blocks_cap, groups));
  const uint32_t iters = (groups + blocks - 1u) / blocks;
  hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(256), 0, ctx->stream, pk, make_float4(norm[0], norm[1], norm[2], norm[3]),
                     rays, labels, n, iters, 1.f / (float)n, partial, loss_part);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  hipLaunchKernelGGL(k_update, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, partial, blocks, loss_part,
                     params, m, v, grad_out, loss_out, 1.f / (float)n, lr, c1, c2, update ? 1 : 0);
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int pack(Ctx* ctx, const float* params, void* packed) {
  hipLaunchKernelGGL(k_pack, dim3((kPackThreads + 255) / 256), dim3(256), 0, ctx->stream, params,
                     reinterpret_cast<Packed*>(packed));
  HIP_TRY(hipGetLastError(), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int mesh_norm(const Mesh& m, float norm[4]) {
  float c[3], d2 = 0.f;
  for (int k = 0; k < 3; ++k) {
    c[k] = 0.5f * (m.info.bb_min[k] + m.info.bb_max[k]);
    const float e = m.info.bb_max[k] - m.info.bb_min[k];
    d2 += e * e;
  }
  const float R = 0.5f * std::sqrt(d2);
  if (!(R > 0.f) || !std::isfinite(R)) return J3DG_ERR_INVALID_ARG;
  norm[0] = c[0], norm[1] = c[1], norm[2] = c[2], norm[3] = R;
  return J3DG_OK;
}

int ensure_mesh_mlp(Ctx* ctx, Mesh& m) {
  if (!m.d_mlp) HIP_TRY(hipMalloc(&m.d_mlp, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  if (!m.d_mlp_packed) HIP_TRY(hipMalloc(&m.d_mlp_packed, sizeof(Packed)), J3DG_ERR_OUT_OF_MEMORY);
  return J3DG_OK;
}

}  // namespace
}  // namespace j3d

// ================================================================== C ABI ===

using j3d::Ctx;
using j3d::DevBuf;
using j3d::Mesh;
using j3d::Target;
using namespace j3d::mlp;

extern "C" {

int j3dg_default_train_opts(j3dg_train_opts* o) {
  if (!o) return J3DG_ERR_INVALID_ARG;
  std::memset(o, 0, sizeof(*o));
  o->n_rays = 1ull << 24;
  o->epochs = 4;
  o->batch = 16384;
  o->lr = 2e-3f;
  o->lr_final = 1e-4f;
  o->seed = 1;
  o->rays_per_camera = 1024;
  o->dist_min = 1.2f;
  o->dist_max = 3.5f;
  o->fov_deg = 60.f;
  o->aspect = 16.f / 9.f;
  o->n_validation = 1u << 16;
  return J3DG_OK;
}

int j3dg_train_proposal(j3dg_ctx* hctx, uint32_t slot, const j3dg_train_opts* opts, j3dg_train_report* rep) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  j3dg_train_opts o;
  if (opts)
    o = *opts;
  else
    j3dg_default_train_opts(&o);
  if (o.n_rays == 0 || o.n_rays > (1ull << 28) || o.batch < 128 || o.batch % 128 || o.epochs == 0 ||
      o.rays_per_camera == 0 || !(o.dist_min > 1.0f) || !(o.dist_max >= o.dist_min) || !(o.fov_deg > 0.f) ||
      !(o.fov_deg < 180.f) || !(o.aspect > 0.f) || !(o.lr > 0.f) || o.n_validation > (1u << 24))
    return ctx->fail(J3DG_ERR_INVALID_ARG, "invalid training options");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  Mesh& m = ctx->meshes[slot];
  float norm[4];
  if (j3d::mesh_norm(m, norm)) return ctx->fail(J3DG_ERR_INVALID_ARG, "mesh has an empty bounding box");
  const uint32_t n = (uint32_t)o.n_rays, nv = o.n_validation;
  const uint32_t nt = n + nv;
  auto t0 = std::chrono::steady_clock::now();

  DevBuf rays, hits, labels, counters;
  HIP_CALL(hipMalloc(&rays.p, (size_t)nt * 8 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&hits.p, (size_t)nt * 5 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&labels.p, (size_t)nt * 2 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&counters.p, 64), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMemsetAsync(counters.p, 0, 64, ctx->stream), J3DG_ERR_LAUNCH);
  const float4 n4 = make_float4(norm[0], norm[1], norm[2], norm[3]);
  const float3 half = make_float3(0.5f * (m.info.bb_max[0] - m.info.bb_min[0]), 0.5f * (m.info.bb_max[1] - m.info.bb_min[1]),
                                  0.5f * (m.info.bb_max[2] - m.info.bb_min[2]));
  const float tan_half = std::tan(0.5f * o.fov_deg * 3.14159265f / 180.f);
  // training rays (seed) and validation rays (seed + 1)
  hipLaunchKernelGGL(k_train_rays, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, o.seed, n, n4, half,
                     o.rays_per_camera, o.dist_min, o.dist_max, tan_half, o.aspect, rays.as<float>());
  if (nv)
    hipLaunchKernelGGL(k_train_rays, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, o.seed + 0x1000001ull, nv, n4,
                       half, o.rays_per_camera, o.dist_min, o.dist_max, tan_half, o.aspect,
                       rays.as<float>() + 8 * (size_t)n);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  // labels from the plain traversal of this mesh in its own (object) space
  {
    j3d::DevObject ob;
    std::memset(&ob, 0, sizeof(ob));
    ob.nodes = m.d_nodes;
    ob.tris = m.d_tris;
    ob.verts = m.d_verts;
    ob.idx = m.d_idx;
    ob.normals = m.d_normals;
    ob.fin = m.d_fin;
    for (int i = 0; i < 16; ++i) ob.cs[i] = ob.inv_cs[i] = (i % 5 == 0) ? 1.f : 0.f;
    ob.db_id = m.db_id;
    ob.width = m.info.width;
    int rc = ctx->upload_object_table(&ob, 1);
    if (rc) return rc;
    j3d::RayQuery q;
    std::memset(&q, 0, sizeof(q));
    q.rays = rays.as<float>();
    q.obj = hits.as<uint32_t>();
    q.tri = q.obj + nt;
    q.t = reinterpret_cast<float*>(q.tri + nt);
    q.u = q.t + nt;
    q.v = q.u + nt;
    j3d::FrameParams fp;
    std::memset(&fp, 0, sizeof(fp));
    rc = ctx->reset_counters();
    if (rc) return rc;
    rc = ctx->launch_trace_kernel(j3d::kModeRays, fp, q, nt, 1, m.max_stack, (int)m.info.width, ctx->d_queue);
    if (rc) return rc;
    hipLaunchKernelGGL(k_train_labels, dim3((nt + 255) / 256), dim3(256), 0, ctx->stream, rays.as<float>(), q.obj, q.t,
                       nt, n4, labels.as<float>(), counters.as<uint32_t>());
    HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
    ctx->obj_cache.clear();
  }
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  uint32_t n_hits_all = 0;
  HIP_CALL(hipMemcpy(&n_hits_all, counters.p, 4, hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  auto t1 = std::chrono::steady_clock::now();

  // mean normalised hit distance of the training set (host copy of the labels)
  std::vector<float> lab((size_t)n * 2);
  HIP_CALL(hipMemcpy(lab.data(), labels.p, lab.size() * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  double s_sum = 0.0;
  uint64_t s_cnt = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (lab[2 * (size_t)i + 1] > 0.5f) s_sum += lab[2 * (size_t)i], ++s_cnt;
  const float s0 = s_cnt ? (float)(s_sum / (double)s_cnt) : 1.f;

  std::vector<float> p0;
  if (o.warm_start && m.has_proposal) {
    p0.resize(kParams);
    HIP_CALL(hipMemcpy(p0.data(), m.d_mlp, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    if (std::memcmp(norm, m.mlp_norm, sizeof(norm)) != 0)
      return ctx->fail(J3DG_ERR_INVALID_ARG, "warm start: the mesh's box changed since the network was trained");
  } else {
    j3d::init_params(o.seed, s0, p0);
  }
  const uint32_t waves_cap = std::max(1u, (uint32_t)ctx->n_cus);  // gradient blocks
  const uint64_t steps_per_epoch = (n + o.batch - 1) / o.batch;
  const uint64_t steps = steps_per_epoch * o.epochs;
  DevBuf params, mom, vel, partial, lpart, losses, packed;
  HIP_CALL(hipMalloc(&params.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&mom.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&vel.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&partial.p, (size_t)waves_cap * kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&lpart.p, waves_cap * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&losses.p, steps * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&packed.p, sizeof(Packed)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMemcpy(params.p, p0.data(), kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemsetAsync(mom.p, 0, kParams * sizeof(float), ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemsetAsync(vel.p, 0, kParams * sizeof(float), ctx->stream), J3DG_ERR_LAUNCH);
  double b1t = 1.0, b2t = 1.0;
  for (uint64_t st = 0; st < steps; ++st) {
    const uint32_t off = (uint32_t)((st % steps_per_epoch) * o.batch);
    const uint32_t nb = std::min<uint32_t>(o.batch, n - off);
    int rc = j3d::pack(ctx, params.as<float>(), packed.p);
    if (rc) return rc;
    b1t *= 0.9, b2t *= 0.999;
    const double frac = steps > 1 ? (double)st / (double)(steps - 1) : 1.0;
    const float lr = (float)(o.lr_final + 0.5 * (o.lr - o.lr_final) * (1.0 + std::cos(3.14159265358979 * frac)));
    rc = j3d::grad_pass(ctx, packed.as<Packed>(), norm, rays.as<float>() + 8 * (size_t)off,
                        labels.as<float>() + 2 * (size_t)off, nb, partial.as<float>(), lpart.as<float>(), waves_cap,
                        params.as<float>(), mom.as<float>(), vel.as<float>(), nullptr, losses.as<float>() + st, lr,
                        (float)(1.0 / (1.0 - b1t)), (float)(1.0 / (1.0 - b2t)), true);
    if (rc) return rc;
  }
  int rc = j3d::pack(ctx, params.as<float>(), packed.p);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  auto t2 = std::chrono::steady_clock::now();

  // store in the mesh
  rc = j3d::ensure_mesh_mlp(ctx, m);
  if (rc) return rc;
  HIP_CALL(hipMemcpy(m.d_mlp, params.p, kParams * sizeof(float), hipMemcpyDeviceToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(m.d_mlp_packed, packed.p, sizeof(Packed), hipMemcpyDeviceToDevice), J3DG_ERR_LAUNCH);
  std::memcpy(m.mlp_norm, norm, sizeof(norm));
  m.has_proposal = true;
  ctx->obj_cache.clear();

  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    std::vector<float> ls(steps);
    HIP_CALL(hipMemcpy(ls.data(), losses.p, steps * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
    const uint64_t k = std::max<uint64_t>(1, steps / 100);
    double a = 0.0, b = 0.0;
    for (uint64_t i = 0; i < k; ++i) a += ls[i], b += ls[steps - 1 - i];
    rep->loss_first = (float)(a / (double)k);
    rep->loss_last = (float)(b / (double)k);
    rep->steps = steps;
    rep->hit_fraction = (float)((double)n_hits_all / (double)nt);
    rep->label_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
    rep->train_ms = (float)std::chrono::duration<double, std::milli>(t2 - t1).count();
    rep->norm[0] = norm[0], rep->norm[1] = norm[1], rep->norm[2] = norm[2], rep->norm[3] = norm[3];
    if (nv) {
      // held-out rays: distance error and the fraction a bounded pass would resolve
      DevBuf vout;
      HIP_CALL(hipMalloc(&vout.p, (size_t)nv * sizeof(float2)), J3DG_ERR_OUT_OF_MEMORY);
      j3d::FrameParams fp;
      std::memset(&fp, 0, sizeof(fp));
      hipLaunchKernelGGL(k_infer<true>, dim3(j3d::infer_grid(ctx->n_cus, nv)), dim3(256), 0, ctx->stream, fp,
                         (const j3d::DevObject*)nullptr, 0u, packed.as<Packed>(), n4,
                         rays.as<float>() + 8 * (size_t)n, nv, vout.as<float2>());
      HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
      std::vector<float2> y(nv);
      std::vector<float> vl((size_t)nv * 2), vr((size_t)nv * 8), vt(nv);
      HIP_CALL(hipMemcpy(y.data(), vout.p, (size_t)nv * sizeof(float2), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(vl.data(), labels.as<float>() + 2 * (size_t)n, vl.size() * sizeof(float),
                         hipMemcpyDeviceToHost),
               J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(vr.data(), rays.as<float>() + 8 * (size_t)n, vr.size() * sizeof(float),
                         hipMemcpyDeviceToHost),
               J3DG_ERR_LAUNCH);
      HIP_CALL(hipMemcpy(vt.data(), reinterpret_cast<float*>(hits.as<uint32_t>() + 2 * (size_t)nt) + n,
                         (size_t)nv * sizeof(float), hipMemcpyDeviceToHost),
               J3DG_ERR_LAUNCH);
      double err = 0.0;
      uint64_t hits_v = 0, correct = 0, within[4] = {0, 0, 0, 0};
      const float margins[4] = {0.f, 0.01f, 0.03f, 0.1f};
      for (uint32_t i = 0; i < nv; ++i) {
        const bool hit = vl[2 * (size_t)i + 1] > 0.5f;
        const bool pred = y[i].y > 0.f;
        correct += hit == pred;
        if (hit) {
          ++hits_v;
          err += std::fabs(y[i].x - vl[2 * (size_t)i]);
          const float* rr = &vr[8 * (size_t)i];
          float x[6], a;
          const float dl = ray_input(norm, rr[0], rr[1], rr[2], rr[3], rr[4], rr[5], x, a);
          const float tp = (y[i].x * norm[3] + a) / dl;
          for (int q = 0; q < 4; ++q)
            if (pred && tp * (1.f + margins[q]) >= vt[i]) ++within[q];
        }
      }
      rep->val_hit_accuracy = (float)((double)correct / (double)nv);
      rep->val_mean_abs_error = hits_v ? (float)(err / (double)hits_v) : 0.f;
      for (int q = 0; q < 4; ++q) rep->val_resolved[q] = hits_v ? (float)((double)within[q] / (double)hits_v) : 0.f;
    }
  }
  return J3DG_OK;
}

int j3dg_get_proposal(j3dg_ctx* hctx, uint32_t slot, float* params, float norm[4]) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  Mesh& m = ctx->meshes[slot];
  if (!m.has_proposal) return ctx->fail(J3DG_ERR_NOT_FOUND, "mesh has no proposal network");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  if (params) HIP_CALL(hipMemcpy(params, m.d_mlp, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  if (norm) std::memcpy(norm, m.mlp_norm, 4 * sizeof(float));
  return J3DG_OK;
}

int j3dg_set_proposal(j3dg_ctx* hctx, uint32_t slot, const float* params, const float norm[4]) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (slot >= ctx->meshes.size() || !ctx->meshes[slot].live) return ctx->fail(J3DG_ERR_NOT_FOUND, "unknown mesh slot");
  Mesh& m = ctx->meshes[slot];
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  if (!params) {
    m.has_proposal = false;
    ctx->obj_cache.clear();
    return J3DG_OK;
  }
  if (!norm || !(norm[3] > 0.f)) return ctx->fail(J3DG_ERR_INVALID_ARG, "invalid normalisation");
  int rc = j3d::ensure_mesh_mlp(ctx, m);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(m.d_mlp, params, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  rc = j3d::pack(ctx, m.d_mlp, m.d_mlp_packed);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  std::memcpy(m.mlp_norm, norm, 4 * sizeof(float));
  m.has_proposal = true;
  ctx->obj_cache.clear();
  return J3DG_OK;
}

int j3dg_mlp_forward(j3dg_ctx* hctx, const float* params, const float norm[4], const float* rays, uint64_t n,
                     float* out) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (!params || !norm || (n && (!rays || !out))) return ctx->fail(J3DG_ERR_INVALID_ARG, "null argument");
  if (n > (1ull << 28)) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many rays");
  if (n == 0) return J3DG_OK;
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  DevBuf dp, pk, dr, dout;
  HIP_CALL(hipMalloc(&dp.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&pk.p, sizeof(Packed)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dr.p, n * 8 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dout.p, n * sizeof(float2)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMemcpy(dp.p, params, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(dr.p, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  int rc = j3d::pack(ctx, dp.as<float>(), pk.p);
  if (rc) return rc;
  j3d::FrameParams fp;
  std::memset(&fp, 0, sizeof(fp));
  hipLaunchKernelGGL(k_infer<true>, dim3(j3d::infer_grid(ctx->n_cus, (uint32_t)n)), dim3(256), 0, ctx->stream, fp,
                     (const j3d::DevObject*)nullptr, 0u, pk.as<Packed>(), make_float4(norm[0], norm[1], norm[2], norm[3]),
                     dr.as<float>(), (uint32_t)n, dout.as<float2>());
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(out, dout.p, n * sizeof(float2), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_mlp_gradients(j3dg_ctx* hctx, const float* params, const float norm[4], const float* rays,
                       const float* labels, uint64_t n, float* grads, float* loss) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (!params || !norm || !grads || !n || !rays || !labels) return ctx->fail(J3DG_ERR_INVALID_ARG, "null argument");
  if (n > (1ull << 24)) return ctx->fail(J3DG_ERR_INVALID_ARG, "too many rays");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  const uint32_t waves_cap = std::max(1u, (uint32_t)ctx->n_cus);  // gradient blocks
  DevBuf dp, pk, dr, dl, part, lp, dg, dloss;
  HIP_CALL(hipMalloc(&dp.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&pk.p, sizeof(Packed)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dr.p, n * 8 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dl.p, n * 2 * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&part.p, (size_t)waves_cap * kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&lp.p, waves_cap * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dg.p, kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMalloc(&dloss.p, sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  HIP_CALL(hipMemcpy(dp.p, params, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(dr.p, rays, n * 8 * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(dl.p, labels, n * 2 * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  int rc = j3d::pack(ctx, dp.as<float>(), pk.p);
  if (rc) return rc;
  rc = j3d::grad_pass(ctx, pk.as<Packed>(), norm, dr.as<float>(), dl.as<float>(), (uint32_t)n, part.as<float>(),
                      lp.as<float>(), waves_cap, nullptr, nullptr, nullptr, dg.as<float>(), dloss.as<float>(), 0.f, 1.f,
                      1.f, false);
  if (rc) return rc;
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(grads, dg.p, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  if (loss) HIP_CALL(hipMemcpy(loss, dloss.p, sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

int j3dg_mlp_adam_step(j3dg_ctx* hctx, float* params, float* m, float* v, const float* grads, uint32_t step,
                       float lr) {
  Ctx* ctx = reinterpret_cast<Ctx*>(hctx);
  CHECK_CTX();
  if (!params || !m || !v || !grads || step == 0) return ctx->fail(J3DG_ERR_INVALID_ARG, "invalid argument");
  HIP_CALL(hipSetDevice(ctx->device), J3DG_ERR_NO_DEVICE);
  DevBuf buf;
  HIP_CALL(hipMalloc(&buf.p, 4 * kParams * sizeof(float)), J3DG_ERR_OUT_OF_MEMORY);
  float* d = buf.as<float>();
  HIP_CALL(hipMemcpy(d, params, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(d + kParams, m, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(d + 2 * kParams, v, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(d + 3 * kParams, grads, kParams * sizeof(float), hipMemcpyHostToDevice), J3DG_ERR_LAUNCH);
  const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)step)));
  hipLaunchKernelGGL(k_update, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, d + 3 * kParams, 1u,
                     (const float*)nullptr, d, d + kParams, d + 2 * kParams, (float*)nullptr, (float*)nullptr, 1.f, lr,
                     c1, c2, 1);
  HIP_CALL(hipGetLastError(), J3DG_ERR_LAUNCH);
  HIP_CALL(hipStreamSynchronize(ctx->stream), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(params, d, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(m, d + kParams, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  HIP_CALL(hipMemcpy(v, d + 2 * kParams, kParams * sizeof(float), hipMemcpyDeviceToHost), J3DG_ERR_LAUNCH);
  return J3DG_OK;
}

}  // extern "C"
