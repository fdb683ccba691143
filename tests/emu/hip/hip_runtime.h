// A CPU stand-in for the HIP runtime and device built-ins, so that the
// library's kernels and host code (j3d_amd/csrc/j3dg.hip) compile for x86 and
// run on the CPU with their memory accesses checked.  Test infrastructure only
// (tests/test_emu_batch.py builds it); nothing in the product includes it.
//
// Execution model.  A launch runs its blocks one after another on one OS
// thread.  Every thread of a block is a fiber (ucontext) with its own stack;
// the fibers of a 64-lane wave run until each reaches a cross-lane operation
// (ballot, shuffle, readfirstlane, any) or exits.  When every live lane of a
// wave is waiting, one group of lanes waiting at the same operation is
// resolved together, as the hardware executes it for the lanes its exec mask
// holds: the group whose call stack is deepest first (lanes inside a
// divergent branch or a callee run before the lanes waiting where control
// flow joins again).  Waves of a block interleave round robin; __syncthreads
// releases when every live thread of the block waits there.
//
// Memory.  hipMalloc places every allocation at the end of its own 256 MiB
// slot of reserved address space whose remainder is unmapped, so a read or
// write past an allocation's end (or far before its start) faults at once.
// The fault handler names the allocation nearest to the address, the kernel
// and the lane, and exits with status 86: the CPU counterpart of the GPU's
// "illegal memory access".
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

// ---- qualifiers ------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define __shared__ static

// ---- vector types ----------------------------------------------------------
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct uint3 { uint32_t x, y, z; };
struct uint4 { uint32_t x, y, z, w; };
struct int2 { int x, y; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline uint2 make_uint2(uint32_t a, uint32_t b) { return {a, b}; }
struct dim3 {
  uint32_t x = 1, y = 1, z = 1;
  dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};

// ---- bit casts and lane-local integer built-ins -----------------------------
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int __float_as_int(float f) { int u; memcpy(&u, &f, 4); return u; }
inline float __int_as_float(int u) { float f; memcpy(&f, &u, 4); return f; }
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int __clzll(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
inline int __ffs(uint32_t x) { return __builtin_ffs((int)x); }
inline int __ffsll(uint64_t x) { return __builtin_ffsll((long long)x); }
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
inline float __log2f(float x) { return std::log2(x); }
inline uint64_t wall_clock64() { return 0; }
inline uint64_t clock64() { return 0; }
inline int __smid() { return 0; }

inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline uint64_t min(uint64_t a, uint64_t b) { return a < b ? a : b; }
inline uint64_t max(uint64_t a, uint64_t b) { return a > b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

// ---- atomics: one OS thread runs every fiber, so a plain read-modify-write is atomic
inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o + v; return o; }
inline int atomicAdd(int* p, int v) { int o = *p; *p = o + v; return o; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p; *p = o + v; return o;
}
inline uint32_t atomicOr(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o | v; return o; }
inline uint32_t atomicMax(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o > v ? o : v; return o; }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p; *p = o > v ? o : v; return o;
}
inline uint32_t atomicMin(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = o < v ? o : v; return o; }
inline uint32_t atomicExch(uint32_t* p, uint32_t v) { uint32_t o = *p; *p = v; return o; }
inline uint32_t atomicCAS(uint32_t* p, uint32_t c, uint32_t v) { uint32_t o = *p; if (o == c) *p = v; return o; }

// ---- the SIMT emulator ----------------------------------------------------------
namespace emu {

enum OpKind : int { kBallot = 1, kAny = 2, kShfl = 3, kShflXor = 4, kFirstLane = 5, kBarrier = 6 };

// One cross-lane operation of the calling lane: blocks until its group resolves.
uint64_t wave_op(int kind, uint64_t value, int arg, const void* site, const void* frame);
dim3 thread_idx();
dim3 block_idx();
extern dim3 g_block_dim, g_grid_dim;
void* dyn_lds();
// A fenced block of device memory for a file-scope __device__ array.
void* device_array(size_t bytes);
// Whether [p, p + bytes) lies inside one live device allocation (a debugging aid).
bool live(const void* p, size_t bytes);
void set_kernel_name(const char* name);

// Runs every thread of every block of a launch: body() is the kernel call.
void run_grid(const char* name, dim3 grid, dim3 block, size_t lds, const std::function<void()>& body);

// hipLaunchKernel: kernels are found by address in a registry that the
// (patched) source fills when it takes a kernel's address.
using Thunk = std::function<std::function<void()>(void**)>;
void register_kernel(const void* fn, const char* name, Thunk t);
bool lookup_kernel(const void* fn, const char** name, Thunk* t);

template <class T>
T read_arg(void* p) {
  return *static_cast<T*>(p);
}
template <class... P, size_t... I>
std::function<void()> bind_args(void (*k)(P...), void** a, std::index_sequence<I...>) {
  auto tup = std::make_tuple(read_arg<typename std::decay<P>::type>(a[I])...);
  return [k, tup]() { std::apply(k, tup); };
}
template <class... P>
void* reg(void (*k)(P...), const char* name) {
  register_kernel((const void*)k, name,
                  [k](void** a) { return bind_args(k, a, std::index_sequence_for<P...>{}); });
  return (void*)k;
}

// Stream capture: launches made while capturing are recorded, not run.
bool capturing();
void record(const char* name, dim3 grid, dim3 block, size_t lds, std::function<void()> body);

template <class... P, class... A>
void launch(const char* name, void (*k)(P...), dim3 grid, dim3 block, size_t lds, A&&... a) {
  auto tup = std::make_tuple(static_cast<typename std::decay<P>::type>(a)...);
  std::function<void()> body = [k, tup]() { std::apply(k, tup); };
  if (capturing())
    record(name, grid, block, lds, body);
  else
    run_grid(name, grid, block, lds, body);
}

template <class T>
inline T shfl_impl(int kind, T v, int arg, const void* site, const void* frame) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit shuffles");
  uint64_t bits = 0;
  memcpy(&bits, &v, sizeof(T));
  uint64_t r = wave_op(kind, bits, arg, site, frame);
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}

}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::block_idx())
#define blockDim (::emu::g_block_dim)
#define gridDim (::emu::g_grid_dim)

#define EMU_SITE_ static const char emu_site_ = 0
#define __ballot(p) \
  ({ EMU_SITE_; (uint64_t)::emu::wave_op(::emu::kBallot, (p) ? 1u : 0u, 0, &emu_site_, __builtin_frame_address(0)); })
#define __any(p) \
  ({ EMU_SITE_; (int)::emu::wave_op(::emu::kAny, (p) ? 1u : 0u, 0, &emu_site_, __builtin_frame_address(0)); })
#define __shfl(v, src, ...) \
  ({ EMU_SITE_; ::emu::shfl_impl(::emu::kShfl, (v), (int)(src), &emu_site_, __builtin_frame_address(0)); })
#define __shfl_xor(v, m, ...) \
  ({ EMU_SITE_; ::emu::shfl_impl(::emu::kShflXor, (v), (int)(m), &emu_site_, __builtin_frame_address(0)); })
#define __builtin_amdgcn_readfirstlane(v) \
  ({ EMU_SITE_; ::emu::shfl_impl(::emu::kFirstLane, (v), 0, &emu_site_, __builtin_frame_address(0)); })
// The wave's lanes meet here (lockstep: every lane's earlier memory accesses
// happen before any lane's later ones), e.g. a read and a write of one shared
// LDS slot by the lanes of a group.
#define __builtin_amdgcn_wave_barrier() \
  do { EMU_SITE_; ::emu::wave_op(::emu::kAny, 0, 0, &emu_site_, __builtin_frame_address(0)); } while (0)
#define __syncthreads() \
  do { EMU_SITE_; ::emu::wave_op(::emu::kBarrier, 0, 0, &emu_site_, __builtin_frame_address(0)); } while (0)
inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t m, uint32_t acc) {
  const uint32_t l = ::emu::thread_idx().x & 63u;
  return acc + (uint32_t)__builtin_popcount(l >= 32 ? m : (m & ((1u << l) - 1u)));
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t m, uint32_t acc) {
  const uint32_t l = ::emu::thread_idx().x & 63u;
  return acc + (l <= 32 ? 0u : (uint32_t)__builtin_popcount(m & ((1u << (l - 32)) - 1u)));
}

// ---- runtime API -------------------------------------------------------------
typedef int hipError_t;
enum : int { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph* hipGraphExec_t;
typedef struct emu_graph_node* hipGraphNode_t;
enum { hipStreamNonBlocking = 1, hipStreamDefault = 0, hipEventDisableTiming = 2, hipEventDefault = 0,
       hipHostMallocDefault = 0 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
struct hipDeviceProp_t {
  int multiProcessorCount;
  char gcnArchName[256];
  size_t totalGlobalMem;
  char name[256];
};

const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int* d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void** p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags = 0);
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t bytes, unsigned flags = 0) {
  return hipHostMalloc(reinterpret_cast<void**>(p), bytes, flags);
}
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemcpy2D(void* d, size_t dpitch, const void* s, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind k);
hipError_t hipMemset(void* d, int v, size_t n);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* err, char* log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGetSymbolAddress(void** p, const void* sym);
#define HIP_SYMBOL(x) ((const void*)&(x))
hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t s);
#define hipLaunchKernelGGL(k, grid, block, lds, stream, ...) \
  ::emu::launch(#k, k, dim3(grid), dim3(block), (size_t)(lds), __VA_ARGS__)
