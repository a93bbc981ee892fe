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
// Streams are ignored (everything runs in call order) and atomics are plain
// read-modify-writes (fibers switch only at barriers).  Not part of the
// product: the product is the HIP build of the same source.
#pragma once

#include <setjmp.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

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
#define __shared__ static
#define __launch_bounds__(...)

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
  for (;;) {
    for (int t = 0; t < nthreads; t++) {
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

template <class T>
inline T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
template <class T>
inline T atomicMin(T *p, T v) { T o = *p; if (v < o) *p = v; return o; }
template <class T>
inline T atomicExch(T *p, T v) { T o = *p; *p = v; return o; }

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
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t = nullptr) { return hipMemset(d, v, n); }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "hipcpu error"; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }

struct hipcpu_event { std::chrono::steady_clock::time_point t; };
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new hipcpu_event(); return hipSuccess; }
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
