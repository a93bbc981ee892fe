// Test infrastructure: CPU execution of the library's own HIP source.
//
// Compiling enzo-dev_amd/csrc/enzo_ppm.hip with g++ and this directory first
// on the include path gives a CPU build of the whole library -- host
// orchestration and kernels -- for checks that the block emulator cannot
// make: the library's own buffer sizes, launch shapes and kernel loops.
//  * Each work-item of a block runs as a fiber; __syncthreads and friends
//    switch fibers, and a block whose work-items do not all reach the same
//    barrier aborts (on the GPU that is a hang).
//  * Every hipMalloc buffer is placed against an inaccessible guard page, by
//    default right after its last byte (HIPCPU_GUARD=low: right before its
//    first), so an access out of bounds stops the process at once.
//  * Every __shared__ array is filled with a NaN pattern (0xff bytes) when a
//    workgroup starts, as LDS holds whatever the previous occupant left on
//    the GPU; with HIPCPU_LDS_CHECK=1 a read of an LDS cell through the tile
//    accessors that the current workgroup has not written stops the process.
//  * HIPCPU_ORDER=rev runs the work-items of every barrier interval in
//    reverse order, HIPCPU_ORDER=shuf:<seed> in a fresh random order per
//    interval; results must not depend on it (no ordering inside a phase).
// Streams are ignored (everything runs in call order) and atomics are plain
// read-modify-writes (fibers switch only at barriers).  Not part of the
// product: the product is the HIP build of the same source.
#pragma once

#include <execinfo.h>
#include <setjmp.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <vector>

#define HIPCPU_RUNTIME 1
#define __global__
#define __device__
#define __host__
// every __shared__ array in one section, so that a workgroup start can poison
// all of them (the linker defines the section's bounds)
#define __shared__ static __attribute__((section("hipcpu_lds")))
extern "C" char __start_hipcpu_lds[], __stop_hipcpu_lds[];
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorMemoryAllocation = 2, hipErrorLaunchFailure = 3 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
typedef void *hipStream_t;
typedef struct hipcpu_event *hipEvent_t;

namespace hipcpu {

inline dim3 g_thread, g_block, g_bdim, g_gdim;

struct Fiber {
  ucontext_t ctx;   // first entry
  jmp_buf jb;       // later switches (_setjmp/_longjmp keep the signal mask alone)
  bool started = false;
  char *stack = nullptr;
  int state = 0;  // 0 runnable, 1 at a barrier, 2 finished
  long long val = 0;
};

inline jmp_buf g_sched;
inline std::vector<Fiber> g_fib;
inline int g_cur = 0;
inline long long g_bar_count = 0;              // nonzero values at the last barrier
inline std::vector<long long> g_slots;          // values deposited for shuffles
inline std::function<void()> g_body;
constexpr size_t kStack = 256 * 1024;

// LDS poisoning and read-before-write tracking.  Static __shared__ locals of
// non-template functions sit in the section hipcpu_lds (poisoned at every
// workgroup start); g++ ignores the section attribute of template statics,
// so the tile kernels announce their arrays with PPM_LDS_ENTER, which
// poisons an array the first time a work-item of a workgroup reaches it and
// keeps a per-double record of the workgroup that last wrote it.
inline int g_lds_check = -1;      // HIPCPU_LDS_CHECK
inline int g_order = -1;          // HIPCPU_ORDER: 0 forward, 1 reverse, 2 shuffled
inline unsigned long long g_rng = 0;
inline uint32_t g_epoch = 0;      // workgroups started
inline unsigned long long g_lds_reads = 0, g_lds_bad = 0;
struct LdsRegion {
  size_t bytes = 0;
  uint32_t poisoned = 0;          // epoch of the last poisoning
  std::vector<uint32_t> wrote;    // epoch of the last write, per double
};
inline std::map<const char *, LdsRegion> g_regions;
inline const char *g_last_base = nullptr;
inline LdsRegion *g_last_region = nullptr;

// read at every launch, so that one process can switch modes between calls
inline void lds_env() {
  const char *c = getenv("HIPCPU_LDS_CHECK");
  g_lds_check = (c && c[0] == '1') ? 1 : 0;
  const char *o = getenv("HIPCPU_ORDER");
  g_order = 0;
  if (o && !strncmp(o, "rev", 3)) g_order = 1;
  if (o && !strncmp(o, "shuf", 4)) {
    g_order = 2;
    const char *col = strchr(o, ':');
    const unsigned long long seed = col ? strtoull(col + 1, nullptr, 10) * 2654435761ull + 1 : 1;
    static unsigned long long last_seed = 0;
    if (seed != last_seed) g_rng = last_seed = seed;
  }
}

inline unsigned long long next_rand() {  // splitmix64
  unsigned long long z = (g_rng += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// the announced region holding p, and p's double index in it
inline LdsRegion &lds_region(const void *p, size_t &k) {
  const char *c = (const char *)p;
  if (!(g_last_region && c >= g_last_base && c < g_last_base + g_last_region->bytes)) {
    auto it = g_regions.upper_bound(c);
    if (it == g_regions.begin() || c >= std::prev(it)->first + std::prev(it)->second.bytes) {
      fprintf(stderr, "hipcpu: tile LDS accessor outside every announced LDS array\n");
      abort();
    }
    --it;
    g_last_base = it->first;
    g_last_region = &it->second;
  }
  k = (size_t)(c - g_last_base) / 8;
  return *g_last_region;
}

inline void fiber_entry() {
  g_body();
  g_fib[g_cur].state = 2;
  _longjmp(g_sched, 1);
}

// every work-item of the block waits here; returns how many passed v != 0
inline long long barrier(long long v) {
  Fiber &f = g_fib[g_cur];
  f.state = 1;
  f.val = v;
  if (!_setjmp(f.jb)) _longjmp(g_sched, 1);
  return g_bar_count;
}

inline void run_block(int nthreads) {
  if ((int)g_fib.size() < nthreads) g_fib.resize(nthreads);
  for (int t = 0; t < nthreads; t++) {
    Fiber &f = g_fib[t];
    if (!f.stack) f.stack = (char *)malloc(kStack);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, fiber_entry, 0);
    f.state = 0;
    f.started = false;
  }
  g_slots.assign(nthreads, 0);
  memset(__start_hipcpu_lds, 0xff, (size_t)(__stop_hipcpu_lds - __start_hipcpu_lds));
  g_epoch++;  // a new workgroup: announced arrays are poisoned again on entry
  std::vector<int> order(nthreads);
  for (int t = 0; t < nthreads; t++) order[t] = g_order == 1 ? nthreads - 1 - t : t;
  for (;;) {
    if (g_order == 2)
      for (int t = nthreads - 1; t > 0; t--) std::swap(order[t], order[next_rand() % (t + 1)]);
    for (int ti = 0; ti < nthreads; ti++) {
      const int t = order[ti];
      if (g_fib[t].state == 2) continue;
      g_fib[t].state = 0;
      g_cur = t;
      g_thread = dim3(t % g_bdim.x, (t / g_bdim.x) % g_bdim.y, t / (g_bdim.x * g_bdim.y));
      if (!_setjmp(g_sched)) {
        if (g_fib[t].started) _longjmp(g_fib[t].jb, 1);
        g_fib[t].started = true;
        setcontext(&g_fib[t].ctx);
      }
    }
    int done = 0, waiting = 0;
    long long cnt = 0;
    for (int t = 0; t < nthreads; t++) {
      if (g_fib[t].state == 2) done++;
      else { waiting++; cnt += g_fib[t].val != 0; }
    }
    if (waiting == 0) return;
    if (done) {
      fprintf(stderr, "hipcpu: %d work-items finished while %d wait at a barrier (block %u)\n", done,
              waiting, g_block.x);
      abort();
    }
    g_bar_count = cnt;
  }
}

// stream capture: launches and copies made while capturing are recorded and
// replayed by hipGraphLaunch
inline std::vector<std::function<void()>> *g_capture = nullptr;

template <class K, class... A>
void launch_now(K kernel, dim3 grid, dim3 block, A... args) {
  lds_env();
  g_gdim = grid;
  g_bdim = block;
  const int nt = (int)(block.x * block.y * block.z);
  for (unsigned z = 0; z < grid.z; z++)
    for (unsigned y = 0; y < grid.y; y++)
      for (unsigned x = 0; x < grid.x; x++) {
        g_block = dim3(x, y, z);
        g_body = [&]() { kernel(args...); };
        run_block(nt);
      }
}

template <class K, class... A>
void launch(K kernel, dim3 grid, dim3 block, A... args) {
  if (g_capture) g_capture->push_back([=]() { launch_now(kernel, grid, block, args...); });
  else launch_now(kernel, grid, block, args...);
}

// guarded allocations: base of the mapping and its size, by user pointer
inline std::map<void *, std::pair<void *, size_t>> g_allocs;

inline void *guarded_alloc(size_t n) {
  const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
  const size_t body = ((n + 7) / 8 * 8 + pg - 1) / pg * pg;
  const size_t total = body + 2 * pg;
  char *base = (char *)mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (base == MAP_FAILED) return nullptr;
  mprotect(base, pg, PROT_NONE);
  mprotect(base + pg + body, pg, PROT_NONE);
  const char *g = getenv("HIPCPU_GUARD");
  char *p = (g && g[0] == 'l') ? base + pg : base + pg + body - (n + 7) / 8 * 8;
  memset(p, 0xff, n);  // like device memory: not zero (NaN pattern)
  g_allocs[p] = {base, total};
  return p;
}

}  // namespace hipcpu

// reads of LDS before any write by the workgroup, counted since the library
// was loaded (HIPCPU_LDS_CHECK=1 with HIPCPU_LDS_CONTINUE set: count, not abort)
extern "C" __attribute__((weak)) unsigned long long hipcpu_lds_violations() { return hipcpu::g_lds_bad; }

// hooks of the tile accessors (ppm_sweep.hpp, PPM_LDS_TRACK)
#define PPM_LDS_TRACK 1
inline void ppm_lds_enter(void *base, size_t bytes) {
  hipcpu::LdsRegion &r = hipcpu::g_regions[(const char *)base];
  if (r.bytes == 0) {
    r.bytes = bytes;
    r.wrote.assign(bytes / 8 + 1, 0);
    hipcpu::g_last_region = nullptr;
  }
  if (r.poisoned != hipcpu::g_epoch) {  // first work-item of this workgroup here
    memset(base, 0xff, bytes);
    r.poisoned = hipcpu::g_epoch;
  }
}
inline void ppm_lds_note_write(const void *p) {
  if (hipcpu::g_lds_check <= 0) return;
  size_t k;
  hipcpu::LdsRegion &r = hipcpu::lds_region(p, k);
  r.wrote[k] = hipcpu::g_epoch;
}
inline void ppm_lds_note_read(const void *p) {
  if (hipcpu::g_lds_check <= 0) return;
  hipcpu::g_lds_reads++;
  size_t k;
  hipcpu::LdsRegion &r = hipcpu::lds_region(p, k);
  if (r.wrote[k] != hipcpu::g_epoch) {
    hipcpu::g_lds_bad++;
    if (hipcpu::g_lds_bad <= 20)
      fprintf(stderr, "hipcpu: LDS read before write: double %zu of an array of %zu bytes, block %u, work-item %u\n",
            k, r.bytes, hipcpu::g_block.x, hipcpu::g_thread.x);
    if (getenv("HIPCPU_LDS_TRACE")) {  // call sites, for addr2line
      void *bt[12];
      backtrace_symbols_fd(bt, backtrace(bt, 12), 2);
    }
    if (!getenv("HIPCPU_LDS_CONTINUE")) abort();
  }
}

#define threadIdx hipcpu::g_thread
#define blockIdx hipcpu::g_block
#define blockDim hipcpu::g_bdim
#define gridDim hipcpu::g_gdim

#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) hipcpu::launch(K, dim3(G), dim3(B), __VA_ARGS__)

inline void __syncthreads() { hipcpu::barrier(0); }
inline int __syncthreads_count(int v) { return (int)hipcpu::barrier(v); }
inline void __threadfence() {}

inline double __shfl_xor(double v, int lane_mask, int width = 64) {
  (void)width;
  long long bits;
  memcpy(&bits, &v, 8);
  const int t = hipcpu::g_cur;
  hipcpu::g_slots[t] = bits;
  hipcpu::barrier(0);
  long long o = hipcpu::g_slots[t ^ lane_mask];
  hipcpu::barrier(0);
  double r;
  memcpy(&r, &o, 8);
  return r;
}

// wavefront vote: bit l of the result is the predicate of work-item
// (t & ~63) + l, for the 64 work-items of t's wavefront (called by every
// work-item of the workgroup, as the kernels do)
inline unsigned long long __ballot(int pred) {
  const int t = hipcpu::g_cur;
  hipcpu::g_slots[t] = pred != 0;
  hipcpu::barrier(0);
  unsigned long long m = 0;
  const int w0 = t & ~63;
  for (int l = 0; l < 64 && w0 + l < (int)hipcpu::g_slots.size(); l++)
    if (hipcpu::g_slots[w0 + l]) m |= 1ull << l;
  hipcpu::barrier(0);
  return m;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

template <class T>
inline T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
template <class T>
inline T atomicMin(T *p, T v) { T o = *p; if (v < o) *p = v; return o; }
template <class T>
inline T atomicExch(T *p, T v) { T o = *p; *p = v; return o; }
// work-items run one at a time: a relaxed atomic load is a plain load
#define __HIP_MEMORY_SCOPE_AGENT 3
template <class T>
inline T __hip_atomic_load(T *p, int, int) { return *p; }

inline long long __double_as_longlong(double d) { long long r; memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; memcpy(&r, &l, 8); return r; }

inline hipError_t hipMalloc(void **p, size_t n) {
  *p = hipcpu::guarded_alloc(n ? n : 8);
  return *p ? hipSuccess : hipErrorMemoryAllocation;
}
inline hipError_t hipFree(void *p) {
  auto it = hipcpu::g_allocs.find(p);
  if (it == hipcpu::g_allocs.end()) return p ? hipErrorInvalidValue : hipSuccess;
  munmap(it->second.first, it->second.second);
  hipcpu::g_allocs.erase(it);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t = nullptr) {
  if (hipcpu::g_capture) {
    hipcpu::g_capture->push_back([=]() { memmove(d, s, n); });
    return hipSuccess;
  }
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
// pinned host memory: ordinary memory on the CPU
constexpr unsigned hipHostMallocDefault = 0;
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned = 0) {
  *p = malloc(n ? n : 8);
  return *p ? hipSuccess : hipErrorMemoryAllocation;
}
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "hipcpu error"; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// a __device__ variable is an ordinary host object here: its address is its
// device address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
// one "compute unit": the CPU runs one workgroup at a time
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 1; return hipSuccess; }

struct hipcpu_event { std::chrono::steady_clock::time_point t; };
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new hipcpu_event(); return hipSuccess; }
constexpr unsigned hipEventDisableTiming = 2;
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}

enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
struct hipcpu_graph { std::vector<std::function<void()>> ops; };
typedef hipcpu_graph *hipGraph_t;
typedef hipcpu_graph *hipGraphExec_t;
typedef struct hipcpu_node *hipGraphNode_t;
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = (hipStream_t) new int(0); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete (int *)s; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  hipcpu::g_capture = new std::vector<std::function<void()>>();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
  *g = new hipcpu_graph{std::move(*hipcpu::g_capture)};
  delete hipcpu::g_capture;
  hipcpu::g_capture = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new hipcpu_graph{g->ops};
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &op : e->ops) op();
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
