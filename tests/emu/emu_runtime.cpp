// The CPU stand-in for the HIP runtime declared in tests/emu/hip/hip_runtime.h:
// fibers in lockstep per wave for the kernels, guard-slot allocations for
// device memory, a fault handler that reports the access, stream capture.
// Test infrastructure only.
#include <execinfo.h>
#include <signal.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>

#include "hip/hip_runtime.h"

namespace emu {

dim3 g_block_dim, g_grid_dim;

namespace {

// ---- device memory: one allocation at the end of each 256 MiB slot ----------
constexpr size_t kSlot = size_t(256) << 20;
constexpr size_t kSlots = 2048;
constexpr size_t kPage = 4096;
char* g_arena = nullptr;
struct Alloc {
  bool live = false;
  char* start = nullptr;
  size_t bytes = 0;
  uint64_t serial = 0;
};
// constructed on first use: a file-scope __device__ array is allocated while
// other translation units' statics are still being initialised
std::vector<Alloc>& allocs() {
  static std::vector<Alloc> v;
  return v;
}
uint64_t g_serial = 0;

char* arena() {
  if (!g_arena) {
    void* p = mmap(nullptr, kSlot * kSlots, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) {
      fprintf(stderr, "emu: cannot reserve the device arena\n");
      abort();
    }
    g_arena = static_cast<char*>(p);
    allocs().resize(kSlots);
  }
  return g_arena;
}

// ---- fibers -------------------------------------------------------------------
constexpr size_t kFiberStack = size_t(512) << 10;
struct Lane {
  ucontext_t ctx;
  char* stack = nullptr;
  int state = 0;  // 0 runnable, 1 waiting at an operation, 2 done
  int kind = 0;
  const void* site = nullptr;
  uintptr_t depth = 0;
  uint64_t val = 0;
  int arg = 0;
  uint64_t result = 0;
  dim3 tid;
  dim3 block;
  size_t block_slot = 0;
};
std::vector<Lane> g_lanes;
ucontext_t g_sched;
Lane* g_cur = nullptr;
dim3 g_cur_block;
std::vector<char> g_lds;
size_t g_lds_bytes = 16;
const std::function<void()>* g_body = nullptr;
const char* g_kernel = "";
uint64_t g_inactive_reads = 0;

void trampoline() {
  (*g_body)();
  g_cur->state = 2;
  swapcontext(&g_cur->ctx, &g_sched);
}

// Resolves the deepest group of waiting lanes of wave [w0, w0 + n).  Returns
// false if the wave has no lane waiting at a cross-lane operation.
bool resolve_wave(size_t w0, size_t n) {
  Lane* first = nullptr;
  for (size_t i = w0; i < w0 + n; ++i) {
    Lane& l = g_lanes[i];
    if (l.state != 1 || l.kind == kBarrier) continue;
    if (!first || l.depth > first->depth) first = &l;
  }
  if (!first) return false;
  const void* site = first->site;
  const uintptr_t depth = first->depth;
  uint64_t group = 0, ballot = 0;
  for (size_t i = w0; i < w0 + n; ++i) {
    Lane& l = g_lanes[i];
    if (l.state == 1 && l.kind != kBarrier && l.site == site && l.depth == depth) {
      group |= 1ull << (i - w0);
      if (l.val) ballot |= 1ull << (i - w0);
    }
  }
  size_t lowest = w0 + (size_t)__builtin_ctzll(group);
  for (size_t i = w0; i < w0 + n; ++i) {
    if (!((group >> (i - w0)) & 1ull)) continue;
    Lane& l = g_lanes[i];
    switch (l.kind) {
      case kBallot:
        l.result = ballot;
        break;
      case kAny:
        l.result = ballot != 0;
        break;
      case kFirstLane:
        l.result = g_lanes[lowest].val;
        break;
      case kShfl:
      case kShflXor: {
        const size_t src = l.kind == kShfl ? ((size_t)l.arg & 63u) : (((i - w0) ^ (size_t)l.arg) & 63u);
        if (src < n && ((group >> src) & 1ull)) {
          l.result = g_lanes[w0 + src].val;
        } else {
          ++g_inactive_reads;  // reading a lane outside the exec mask: undefined on the GPU
          l.result = 0xDEADBEEFDEADBEEFull;
        }
        break;
      }
      default:
        break;
    }
    l.state = 0;
  }
  return true;
}

void fault_handler(int sig, siginfo_t* si, void*) {
  const char* a = static_cast<const char*>(si->si_addr);
  char msg[512];
  int k = -1;
  long long off = 0;
  uint64_t serial = 0;
  size_t bytes = 0;
  bool freed = false;
  if (g_arena && a >= g_arena && a < g_arena + kSlot * kSlots) {
    const size_t s = (size_t)(a - g_arena) / kSlot;
    // the allocation this slot holds (or held: a freed one), or the one before it (an access past its end)
    for (long long j = (long long)s; j >= 0 && j >= (long long)s - 1; --j) {
      const Alloc& al = allocs()[(size_t)j];
      if (al.start) {
        k = (int)j;
        off = (long long)(a - al.start);
        serial = al.serial;
        bytes = al.bytes;
        freed = !al.live;
        break;
      }
    }
  }
  int n = snprintf(msg, sizeof(msg),
                   "EMU FAULT: signal %d at %p in kernel %s, block (%u,%u,%u) thread %u: %s allocation #%llu "
                   "(%zu bytes) at offset %lld\n",
                   sig, (const void*)a, g_kernel, g_cur ? g_cur->block.x : 0u, g_cur ? g_cur->block.y : 0u,
                   g_cur ? g_cur->block.z : 0u, g_cur ? g_cur->tid.x : 0u,
                   k < 0 ? "no" : freed ? "a freed" : "outside", (unsigned long long)serial, bytes, off);
  if (n > 0) (void)!write(2, msg, (size_t)std::min(n, (int)sizeof(msg) - 1));
  void* frames[48];
  const int depth = backtrace(frames, 48);
  backtrace_symbols_fd(frames, depth, 2);
  _exit(86);
}

struct Installer {
  Installer() {
    static char altstack[1 << 16];
    stack_t ss;
    ss.ss_sp = altstack;
    ss.ss_size = sizeof(altstack);
    ss.ss_flags = 0;
    sigaltstack(&ss, nullptr);
    struct sigaction sa;
    memset(&sa, 0, sizeof(sa));
    sa.sa_sigaction = fault_handler;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);
  }
} g_installer;

struct Kernel {
  const char* name;
  Thunk thunk;
};
std::map<const void*, Kernel>& registry() {
  static std::map<const void*, Kernel> r;
  return r;
}

bool g_capturing = false;
struct Recorded {
  const char* name;
  dim3 grid, block;
  size_t lds;
  std::function<void()> body;
};
std::vector<Recorded> g_recording;

}  // namespace

uint64_t wave_op(int kind, uint64_t value, int arg, const void* site, const void* frame) {
  Lane& l = *g_cur;
  l.kind = kind;
  l.val = value;
  l.arg = arg;
  l.site = site;
  l.depth = (uintptr_t)(l.stack + kFiberStack) - (uintptr_t)frame;
  l.state = 1;
  swapcontext(&l.ctx, &g_sched);
  return l.result;
}

dim3 thread_idx() { return g_cur->tid; }
dim3 block_idx() { return g_cur->block; }
void* dyn_lds() { return g_lds.data() + g_cur->block_slot * g_lds_bytes; }
void set_kernel_name(const char* name) { g_kernel = name; }

bool live(const void* p, size_t bytes) {
  const char* c = static_cast<const char*>(p);
  if (!g_arena || c < g_arena || c >= g_arena + kSlot * kSlots) return false;
  const Alloc& al = allocs()[(size_t)(c - g_arena) / kSlot];
  return al.live && c >= al.start && c + bytes <= al.start + al.bytes;
}

void* device_array(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) abort();
  return p;
}

// Runs the threads of `blocks` (block indices) together: their waves
// interleave round robin, so a persistent kernel's blocks share its queues as
// resident workgroups do on the GPU.
void run_blocks(const char* name, const std::vector<dim3>& blocks, size_t nt, size_t lds,
                const std::function<void()>& body) {
  const size_t total = nt * blocks.size();
  if (g_lanes.size() < total) g_lanes.resize(total);
  for (size_t i = 0; i < total; ++i)
    if (!g_lanes[i].stack) g_lanes[i].stack = static_cast<char*>(malloc(kFiberStack));
  g_lds.assign(std::max<size_t>(lds, 16) * blocks.size(), (char)0xA5);  // LDS starts undefined
  g_lds_bytes = std::max<size_t>(lds, 16);
  g_kernel = name;
  g_body = &body;
  for (size_t i = 0; i < total; ++i) {
    Lane& l = g_lanes[i];
    l.state = 0;
    l.tid = dim3((uint32_t)(i % nt), 0, 0);
    l.block = blocks[i / nt];
    l.block_slot = i / nt;
    getcontext(&l.ctx);
    l.ctx.uc_stack.ss_sp = l.stack;
    l.ctx.uc_stack.ss_size = kFiberStack;
    l.ctx.uc_link = nullptr;
    makecontext(&l.ctx, trampoline, 0);
  }
  size_t live = total;
  while (live) {
    bool progress = false;
    for (size_t w0 = 0; w0 < total; w0 += 64) {
      const size_t n = std::min<size_t>(64, nt - (w0 % nt));
      for (size_t i = w0; i < w0 + n; ++i) {
        Lane& l = g_lanes[i];
        if (l.state != 0) continue;
        g_cur = &l;
        g_cur_block = l.block;
        swapcontext(&g_sched, &l.ctx);
        progress = true;
        if (l.state == 2) --live;
      }
      if (resolve_wave(w0, n)) progress = true;
    }
    if (live && !progress) {
      // every live thread of some block waits at the barrier: release those blocks
      bool released = false;
      for (size_t b = 0; b < blocks.size(); ++b) {
        size_t blive = 0, at = 0;
        for (size_t i = b * nt; i < (b + 1) * nt; ++i) {
          blive += g_lanes[i].state != 2;
          at += g_lanes[i].state == 1 && g_lanes[i].kind == kBarrier;
        }
        if (blive && at == blive) {
          for (size_t i = b * nt; i < (b + 1) * nt; ++i)
            if (g_lanes[i].state == 1) g_lanes[i].state = 0;
          released = true;
        }
      }
      if (!released) {
        fprintf(stderr, "emu: deadlock in kernel %s\n", name);
        abort();
      }
    }
  }
  g_cur = nullptr;
}

void run_grid(const char* name, dim3 grid, dim3 block, size_t lds, const std::function<void()>& body) {
  const size_t nt = (size_t)block.x * block.y * block.z;
  if (block.y != 1 || block.z != 1 || nt % 64) {
    // the kernels index threads by threadIdx.x only and launch whole waves
    fprintf(stderr, "emu: only 1-D blocks of whole waves are emulated\n");
    abort();
  }
  g_block_dim = block;
  g_grid_dim = grid;
  // The persistent trace kernel's blocks run together (their only static LDS,
  // the object table, holds the same data in every block); other kernels keep
  // per-block static LDS, so their blocks run one after another.
  const bool together = strcmp(name, "k_trace") == 0;
  std::vector<dim3> all;
  for (uint32_t bz = 0; bz < grid.z; ++bz)
    for (uint32_t by = 0; by < grid.y; ++by)
      for (uint32_t bx = 0; bx < grid.x; ++bx) all.push_back(dim3(bx, by, bz));
  if (together) {
    run_blocks(name, all, nt, lds, body);
  } else {
    for (const dim3& b : all) run_blocks(name, std::vector<dim3>{b}, nt, lds, body);
  }
}

void register_kernel(const void* fn, const char* name, Thunk t) { registry()[fn] = Kernel{name, std::move(t)}; }

bool lookup_kernel(const void* fn, const char** name, Thunk* t) {
  auto it = registry().find(fn);
  if (it == registry().end()) return false;
  *name = it->second.name;
  *t = it->second.thunk;
  return true;
}

bool capturing() { return g_capturing; }
void record(const char* name, dim3 grid, dim3 block, size_t lds, std::function<void()> body) {
  g_recording.push_back(Recorded{name, grid, block, lds, std::move(body)});
}

uint64_t inactive_lane_reads() { return g_inactive_reads; }

}  // namespace emu

extern "C" unsigned long long emu_inactive_lane_reads() { return emu::inactive_lane_reads(); }

// ---- runtime API ---------------------------------------------------------------
using namespace emu;

struct emu_stream {
  int dummy;
};
struct emu_event {
  std::chrono::steady_clock::time_point t;
};
struct emu_graph {
  std::vector<Recorded> nodes;
};

const char* hipGetErrorString(hipError_t e) {
  return e == hipSuccess ? "no error" : e == hipErrorOutOfMemory ? "out of memory" : "invalid value";
}
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipPeekAtLastError() { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  memset(p, 0, sizeof(*p));
  // a persistent grid of a few blocks; EMU_CUS=n makes it n times as many
  // (more waves, each holding fewer rays: every wave reaches its tail sooner)
  static const int cus = getenv("EMU_CUS") ? atoi(getenv("EMU_CUS")) : 1;
  p->multiProcessorCount = cus > 0 && cus <= 64 ? cus : 1;
  strcpy(p->gcnArchName, "gfx950-cpu-emulation");
  strcpy(p->name, "CPU emulation");
  p->totalGlobalMem = kSlot * kSlots;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }

hipError_t hipMalloc(void** p, size_t bytes) {
  char* base = arena();
  if (bytes > kSlot - kPage) return hipErrorOutOfMemory;
  // every other slot (a free slot between any two allocations); slots are not
  // reused while unused ones remain, so a stale pointer keeps faulting
  // (EMU_SLOT_REUSE=1: the lowest free slot first, as a caching allocator would)
  static size_t next = 1;
  static const bool reuse = getenv("EMU_SLOT_REUSE") && getenv("EMU_SLOT_REUSE")[0] == '1';
  if (reuse) next = 1;
  for (size_t tries = 0; tries < kSlots / 2; ++tries, next = next + 2 < kSlots ? next + 2 : 1) {
    const size_t s = next;
    if (allocs()[s].live) continue;
    char* slot_end = base + (s + 1) * kSlot;
    char* start = slot_end - ((bytes + 15) & ~size_t(15));
    char* page = reinterpret_cast<char*>(reinterpret_cast<uintptr_t>(start) & ~(uintptr_t)(kPage - 1));
    if (slot_end > page && mprotect(page, (size_t)(slot_end - page), PROT_READ | PROT_WRITE) != 0)
      return hipErrorOutOfMemory;
    memset(page, 0xCD, (size_t)(slot_end - page));  // device memory starts undefined
    allocs()[s] = Alloc{true, start, bytes, ++g_serial};
    next = next + 2 < kSlots ? next + 2 : 1;
    *p = start;
    return hipSuccess;
  }
  return hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  char* c = static_cast<char*>(p);
  const size_t s = (size_t)(c - g_arena) / kSlot;
  if (!g_arena || c < g_arena || s >= kSlots || !allocs()[s].live || allocs()[s].start != c) {
    fprintf(stderr, "emu: hipFree of an address that hipMalloc did not return\n");
    abort();
  }
  char* slot = g_arena + s * kSlot;
  madvise(slot, kSlot, MADV_DONTNEED);
  mprotect(slot, kSlot, PROT_NONE);
  allocs()[s].live = false;
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = malloc(std::max<size_t>(bytes, 1));
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  if (g_capturing) {
    fprintf(stderr, "emu: a copy inside stream capture\n");
    abort();
  }
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemcpy2D(void* d, size_t dpitch, const void* s, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind) {
  for (size_t y = 0; y < height; ++y)
    memcpy(static_cast<char*>(d) + y * dpitch, static_cast<const char*>(s) + y * spitch, width);
  return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n) {
  if (n) memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
  if (g_capturing) {
    fprintf(stderr, "emu: a memset inside stream capture\n");
    abort();
  }
  return hipMemset(d, v, n);
}
hipError_t hipStreamCreate(hipStream_t* s) {
  *s = new emu_stream{0};
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new emu_event{std::chrono::steady_clock::now()};
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  if (e) e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  // EMU_RANDOM_TIME=<seed>: seeded random durations, so that choices made from
  // timings (the autotuner's) take a different path on every seed
  static const char* seed = getenv("EMU_RANDOM_TIME");
  if (seed && seed[0]) {
    static uint64_t x = 0x9E3779B97F4A7C15ull ^ strtoull(seed, nullptr, 10);
    x ^= x << 13, x ^= x >> 7, x ^= x << 17;
    *ms = 0.5f + (float)(x % 1000) / 1000.f;
    return hipSuccess;
  }
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  g_capturing = true;
  g_recording.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  g_capturing = false;
  *g = new emu_graph{std::move(g_recording)};
  g_recording.clear();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = new emu_graph{g->nodes};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (const Recorded& r : e->nodes) run_grid(r.name, r.grid, r.block, r.lds, r.body);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t) {
  const char* name = nullptr;
  Thunk t;
  if (!lookup_kernel(fn, &name, &t)) {
    fprintf(stderr, "emu: hipLaunchKernel of an unregistered kernel\n");
    abort();
  }
  std::function<void()> body = t(args);
  if (g_capturing)
    record(name, grid, block, lds, std::move(body));
  else
    run_grid(name, grid, block, lds, body);
  return hipSuccess;
}
