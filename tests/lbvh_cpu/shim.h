// Host stand-ins for the HIP built-ins the device BVH builder's kernels use,
// so that their source (the lbvh namespace of j3d_amd/csrc/bvh_gpu.hip) can be
// compiled by g++ and run serially, one thread index after another, with
// address and undefined-behaviour sanitizers.  Atomics are plain
// read-modify-writes (one thread at a time); LAUNCH runs every thread of a grid.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "bvh_build.h"
#include "rt_math.h"

#define __global__
#define __device__
#define __host__
#define __restrict__
struct dim3 {
  uint32_t x = 1, y = 1, z = 1;
  dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
static dim3 blockIdx, blockDim, threadIdx, gridDim;
struct float4 { float x, y, z, w; };
struct float3 { float x, y, z; };
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int __clzll(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o + v; return o; }
inline uint32_t atomicMax(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = std::max(o, v); return o; }
inline uint32_t atomicOr(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o | v; return o; }
using std::max;
using std::min;
#define LAUNCH(kern, grid, block, ...)                                     \
  do {                                                                     \
    gridDim = dim3(grid);                                                  \
    blockDim = dim3(block);                                                \
    for (uint32_t b_ = 0; b_ < gridDim.x; ++b_)                            \
      for (uint32_t t_ = 0; t_ < blockDim.x; ++t_) {                       \
        blockIdx = dim3(b_);                                               \
        threadIdx = dim3(t_);                                              \
        kern(__VA_ARGS__);                                                 \
      }                                                                    \
  } while (0)
