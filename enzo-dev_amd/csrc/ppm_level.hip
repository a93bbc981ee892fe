// ppm_level.hip -- batched update of one AMR level (include/enzo_ppm.h,
// hipPPMLevel*): every grid of the level in one device arena, each sweep one
// launch per tile shape over a work list of (grid, tile) items, the fallback
// over a level-wide list, and the launches of a step captured once in a
// hipGraph and replayed.
//
// It computes what a loop of grid::SolvePPM_DE calls over the level's grids
// computes (EvolveLevel.C -> Grid_SolveHydroEquations.C:513-517), grid for
// grid bitwise: every grid keeps its own geometry, coefficient tables,
// segments and flattening rows; the level shares the step's dt.
//
// Layout: arena A and B hold, grid after grid, the NSLOT slots of each grid
// (slot s of grid g at off[g] + s * cells[g]); a sweep reads A and writes B
// and the arenas swap.  Pressure, acceleration, tables, widths, flattening
// rows and dt/dx live in arenas with per-grid offsets too.  SweepParams of
// every (parity, axis, grid) sit in device memory and the kernels read them
// there.  With diffusion the sweep runs slice by slice (the next slice reads
// this one's results), slice k of every grid in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/enzo_ppm.h"
#include "ppm_device.hpp"
#include "ppm_geometry.hpp"
#include "ppm_sweep.hpp"

using namespace ppm;

namespace {

struct Work {
  int grid, tile;
};

// sibling copy box: every slot of grid `dst` at [d0, d0 + n) takes grid
// `src`'s values at [s0, s0 + n) (local, ghost-inclusive indices, x first)
struct CopyBox {
  int dst, src;
  int d0[3], s0[3], n[3];
};

// where a grid's state lives in the arena
struct GridGeom {
  long long off, cells;
  int GD[3];
};

// tile shapes of the level kernels (256 work-items each)
constexpr int kShapes = 3;
constexpr int kShapeTP[kShapes] = {16, 8, 4};
constexpr int kShapeTS[kShapes] = {16, 32, 64};

template <int TP, int TS, int UNI, int LEARN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
void level_sweep_kernel(const SweepParams *__restrict__ params, const Work *__restrict__ work) {
  __shared__ double sm[Cfg<TP, TS>::SMEM];
  PPM_LDS_ENTER(sm);
  DevExec ex;
  const Work w = work[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
  const SweepParams &P = params[w.grid];
  TileCounts tc = sweep_tile<TP, TS, UNI, DevExec, LEARN>(P, w.tile, sm, ex);
  if (threadIdx.x == 0) {
    if (tc.bad || tc.err) flag_group(P, w.tile / P.nseg);
    if (tc.chains) atomicAdd(&P.counters[1], (unsigned long long)tc.chains);
  }
}

// The same sweep as one persistent launch: a few workgroups per CU, each
// pulling work items until the list is done.  The items are split into one
// contiguous range per XCD (workgroup b runs on XCD b % kXCD, each XCD with
// its own L2; with fewer than kXCD workgroups, one range per workgroup), and
// the workgroups of a range take its items in order from a counter of their
// own, queue[x]; queue[kXCD + x] counts the range's workgroups that have
// finished, and the last of them puts both back to 0 for the next launch
// (graph replays included).  Every workgroup leaves the loop once its pull
// passes its range's end, so every item is taken exactly once.
template <int TP, int TS, int UNI, int LEARN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void level_sweep_persistent_kernel(const SweepParams *__restrict__ params,
                                   const Work *__restrict__ work, int nitems, unsigned *queue) {
  __shared__ double sm[Cfg<TP, TS>::SMEM];
  __shared__ int next;
  PPM_LDS_ENTER(sm);
  const int nparts = (int)gridDim.x < kXCD ? (int)gridDim.x : kXCD;
  const int x = (int)(blockIdx.x % nparts);
  const int q = nitems / nparts, rem = nitems % nparts;
  const int lo = x * q + (x < rem ? x : rem), hi = lo + q + (x < rem ? 1 : 0);
  for (;;) {
    if (threadIdx.x == 0) next = lo + (int)atomicAdd(&queue[x], 1u);
    __syncthreads();
    // uniform over the workgroup: the item, and with it the grid's
    // parameters, go to scalar registers as in the one-item kernel
    const int k = PPM_UNIFORM(next);
    __syncthreads();  // `next` is free for the next pull
    if (k >= hi) break;
    const Work w = work[k];
    const SweepParams &P = params[w.grid];
    DevExec ex;
    // the work-item index as a value of this iteration: otherwise the
    // compiler hoists every tid-derived address of the tile out of the loop
    // and keeps them all live (spilling registers)
    PPM_OPAQUE_VGPR(ex.tid);
    TileCounts tc = sweep_tile<TP, TS, UNI, DevExec, LEARN>(P, w.tile, sm, ex);
    if (threadIdx.x == 0) {
      if (tc.bad || tc.err) flag_group(P, w.tile / P.nseg);
      if (tc.chains) atomicAdd(&P.counters[1], (unsigned long long)tc.chains);
    }
  }
  if (threadIdx.x == 0) {
    const unsigned nx = gridDim.x / nparts + ((int)(gridDim.x % nparts) > x ? 1u : 0u);
    __threadfence();
    if (atomicAdd(&queue[kXCD + x], 1u) == nx - 1) {
      queue[x] = 0u;
      queue[kXCD + x] = 0u;
      __threadfence();
    }
  }
}

// fallback over the level's list: entries are level-wide group indices
// (the grid's first group + the group), grp_grid maps them to their grid
template <int LEARN, int TS = kFbTS>
__global__ __launch_bounds__(TS) void level_fallback_kernel(const SweepParams *__restrict__ params,
                                                              const unsigned *list,
                                                              const int *grp_grid,
                                                              const unsigned *grp_off,
                                                              unsigned *count, unsigned *done) {
  __shared__ double sm[NARR_FB * Cfg<1, TS>::LW];
  PPM_LDS_ENTER(sm);
  DevExec ex;
  const unsigned n = *count;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
    const unsigned e = list[k];
    const int g = grp_grid[e];
    const SweepParams &P = params[g];
    if (P.uni[0]) fallback_group<1, LEARN, TS>(P, (int)(e - grp_off[g]), sm, ex);
    else fallback_group<0, LEARN, TS>(P, (int)(e - grp_off[g]), sm, ex);
  }
  fallback_launch_done(count, done);
}

// per-sweep preparation of every grid: one workgroup per grid
__global__ void level_prep_kernel(const FlatParams *__restrict__ F) {
  const FlatParams &f = F[blockIdx.y];
  const long long n = f.N + (f.iflatten ? (long long)(f.i2 - f.i1 + 3) * f.n_outer : 0);
  for (long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x; gid < n;
       gid += (long long)gridDim.x * blockDim.x)
    prep_item(f, gid);
}

// pressure + Courant over chunks of kPresChunk cells: work = (grid, first cell)
// (4 cells per work-item: at 518^3 the pressure pass takes 1.31 ms, the
// expansion terms 3.07 ms and the old-state copy 0.74 ms, against 1.37,
// 3.14 and 0.81 ms with chunks of 4096)
constexpr int kPresChunk = 1024;
__global__ __launch_bounds__(256) void level_pressure_kernel(const PresParams *__restrict__ Q,
                                                             const Work *__restrict__ work) {
  const Work w = work[blockIdx.x];
  const PresParams &q = Q[w.grid];
  double local = 1.0e20;
  const long long lim = (long long)w.tile + kPresChunk;
  const long long end = q.cells < lim ? q.cells : lim;
  for (long long g = (long long)w.tile + threadIdx.x; g < end; g += blockDim.x)
    local = fmn(local, pressure_dt_cell(q, g));
  if (q.want_dt) block_min_dt(local, q.dtmin);
}

// comoving: the slots of the state at the step's start that the expansion
// terms read (expansion_old_slots), kept per chunk
__global__ __launch_bounds__(256) void level_save_old_kernel(const ExpandParams *__restrict__ X,
                                                             const Work *__restrict__ work) {
  const Work w = work[blockIdx.x];
  const ExpandParams &x = X[w.grid];
  const long long lim = (long long)w.tile + kPresChunk;
  const long long end = x.cells < lim ? x.cells : lim;
  int slots[5];
  const int ns = expansion_old_slots(x.idual, x.rank, slots);
  for (int k = 0; k < ns; k++) {
    const long long s = slots[k];
    for (long long g = (long long)w.tile + threadIdx.x; g < end; g += blockDim.x)
      x.old[s * x.cells + g] = x.A[s * x.cells + g];
  }
}

// comoving expansion terms per chunk
__global__ __launch_bounds__(256) void level_expand_kernel(const ExpandParams *__restrict__ X,
                                                           const Work *__restrict__ work) {
  const Work w = work[blockIdx.x];
  const ExpandParams &x = X[w.grid];
  const long long lim = (long long)w.tile + kPresChunk;
  const long long end = x.cells < lim ? x.cells : lim;
  for (long long g = (long long)w.tile + threadIdx.x; g < end; g += blockDim.x) expand_cell(x, g);
}

// sibling copies (grid::CopyZonesFromGrid): one workgroup row per box
__global__ void level_copy_kernel(const CopyBox *__restrict__ boxes,
                                  const GridGeom *__restrict__ geom, double *S, int nslot) {
  const CopyBox &b = boxes[blockIdx.y];
  const GridGeom &gd = geom[b.dst], &gs = geom[b.src];
  const long long per_slot = (long long)b.n[0] * b.n[1] * b.n[2];
  const long long n = per_slot * nslot;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const int slot = (int)(q / per_slot);
    long long r = q % per_slot;
    const int x = (int)(r % b.n[0]);
    r /= b.n[0];
    const int y = (int)(r % b.n[1]);
    const int z = (int)(r / b.n[1]);
    const long long id =
        ((long long)(b.d0[2] + z) * gd.GD[1] + (b.d0[1] + y)) * gd.GD[0] + (b.d0[0] + x);
    const long long is =
        ((long long)(b.s0[2] + z) * gs.GD[1] + (b.s0[1] + y)) * gs.GD[0] + (b.s0[0] + x);
    S[gd.off + (long long)slot * gd.cells + id] = S[gs.off + (long long)slot * gs.cells + is];
  }
}

// periodic ghost fill of every grid along one axis: one workgroup row per grid
__global__ void level_periodic_kernel(const BCParams *__restrict__ B) {
  const BCParams &b = B[blockIdx.y];
  const int n = (int)periodic_items(b);  // below 2^31 (periodic_fits)
  for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += gridDim.x * blockDim.x)
    periodic_item(b, gid);
}

// RiemannSolver HLL / HLLC: the launch's group list onto the fallback list
__global__ void level_fb_fill_kernel(const unsigned *__restrict__ src, unsigned n, unsigned *list,
                                     unsigned *count) {
  for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    list[k] = src[k];
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

// one domain face of every grid in the list (blockIdx.y)
__global__ void level_ext_bc_kernel(const ExtBCParams *__restrict__ B) {
  const ExtBCParams &b = B[blockIdx.y];
  const long long n = ext_bc_items(b);
  for (long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x; gid < n;
       gid += (long long)gridDim.x * blockDim.x)
    ext_bc_item(b, gid);
}

__global__ void level_dt_init_kernel(unsigned long long *dtmin) {
  if (threadIdx.x == 0) *dtmin = (unsigned long long)__double_as_longlong(1.0e20);
}

__global__ void level_dt_final_kernel(const unsigned long long *dtmin, double cfl, double *dt) {
  if (threadIdx.x == 0) *dt = __longlong_as_double((long long)*dtmin) * cfl;
}

__global__ void level_zero_kernel(unsigned long long *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

thread_local char g_level_error[512] = "";

void level_error(const char *what, hipError_t e, int line) {
  snprintf(g_level_error, sizeof(g_level_error), "%s failed: %s (ppm_level.hip:%d)", what,
           hipGetErrorString(e), line);
}

#define LCHK(x)                              \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) {                  \
      level_error(#x, e_, __LINE__);         \
      return ENZO_PPM_ERR_HIP;               \
    }                                        \
  } while (0)

template <class T>
bool dev_alloc(T **p, size_t n) {
  *p = nullptr;
  return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess;
}

template <int TP, int TS>
bool segments_for(int N, int i1, int i2, std::vector<int> &o, std::vector<int> &w) {
  return build_segments<TP, TS>(N, i1, i2, o, w);
}

bool segments_shape(int s, int N, int i1, int i2, std::vector<int> &o, std::vector<int> &w) {
  if (s == 0) return segments_for<16, 16>(N, i1, i2, o, w);
  if (s == 1) return segments_for<8, 32>(N, i1, i2, o, w);
  return segments_for<4, 64>(N, i1, i2, o, w);
}

}  // namespace

struct hipPPMLevel {
  int ngrids = 0, rank = 0, nslot = 0, ncol = 0;
  bool gravity = false, diffusion = false;
  hipPPMParameter para0;            // the level's hydro switches (grid 0's block)
  std::vector<hipPPMParameter> para;
  // per grid
  std::vector<int> GD;              // 3 per grid
  std::vector<long long> cells, off, poff;
  std::vector<long long> tab_off, cw_off;  // 3 per grid (doubles into tabs / cws)
  std::vector<long long> flat_off, qc_off;
  std::vector<int> uni;             // 3 per grid
  std::vector<double> kt;           // 3 * NTAB per grid
  std::vector<int> shape;           // 3 per grid: tile shape index per sweep axis
  // segments per (grid, axis, shape): o then w, in segs
  std::vector<int> hsegs;
  std::vector<long long> seg_off;   // ((g * 3 + d) * kShapes + s) -> offset into segs
  std::vector<int> seg_n;
  std::vector<long long> grp_off;   // fallback group flags per grid
  long long ngroups = 0;
  // device memory
  double *A = nullptr, *B = nullptr, *pres = nullptr, *accel = nullptr;
  double *tabs = nullptr, *cws = nullptr, *flat = nullptr, *qc = nullptr;
  int *segs = nullptr;
  double *dt = nullptr;
  unsigned long long *dtmin = nullptr, *counters = nullptr;
  int *fb_flag = nullptr;
  unsigned *fb_list = nullptr, *fb_state = nullptr;
  int *grp_grid = nullptr;          // level-wide group -> grid
  unsigned *dgrp_off = nullptr;     // grid -> its first level-wide group
  SweepParams *dparams = nullptr;   // [parity 2][axis 3][grid]
  FlatParams *dflat = nullptr;      // [parity 2][axis 3][grid]
  PresParams *dpres = nullptr;      // [parity 2][grid]
  Work *dwork = nullptr;            // all work lists
  // RiemannSolver HLL / HLLC: every pencil group goes to the one-pencil
  // workgroups; per axis and slice, the level-wide group list
  int whole = 0;
  unsigned *dfball = nullptr;
  std::vector<long long> fball_first[3];
  int fb_ts[3] = {kFbTS, kFbTS, kFbTS};  // workgroup width holding every grid's pencils
  Work *dpwork = nullptr;           // pressure chunks
  long long npwork = 0;
  // work lists: per axis, per launch: (shape, uni, first item, count, slice)
  struct Launch {
    int shape, uni;
    long long first, count;
    long long queue;  // its 2 * kXCD counters of the persistent launch, at queues + 2 kXCD queue
  };
  std::vector<Launch> launches[3];
  std::vector<long long> launch_slices[3];  // index of the first launch of each slice
  int prep_blocks[3] = {1, 1, 1};           // workgroups per grid of the prep launch
  std::vector<SweepParams> hparams;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int parity = 0;                   // 0: state in A
  // the pressure array holds the current state's pressure
  // (hipPPMLevelTimeStep computed it and nothing has changed the state
  // since), so the next hipPPMLevelStep need not compute it again
  bool pres_current = false;
  bool use_graph = true;
  // persistent sweep launches (level_sweep_persistent_kernel) instead of one
  // workgroup per work item; persistent_blocks workgroups per launch
  bool persistent = false;
  int persistent_blocks = 1024;
  unsigned *queues = nullptr;
  // the last hipPPMLevelTune's timings, seconds per level step: [class][shape]
  // (NaN: not timed), then the ordinary and the persistent launch
  double tune_times[4 * kShapes + 2];
  std::map<int, hipGraphExec_t> graphs;
  double graph_cfl = -1.0;          // cfl the Courant-step graphs were captured with
  long long fallback_total = 0;
  float *learn = nullptr;           // learned two-shock start, when `learned`
  bool learned = false;
  // comoving coordinates: sweeps on widths a_sweep * CellWidth, Courant pass
  // on CellWidth (cws_dt) with aye, the state at each step's start in `old`
  bool comoving = false, old_valid = false;
  double a_sweep = 1.0, aye = 1.0;
  std::vector<double> hcw;          // unscaled widths, cw_off layout
  double *cws_dt = nullptr;
  double *old = nullptr;            // (S_GE + 1) * cells per grid at ooff
  std::vector<long long> ooff;
  ExpandParams *dexp = nullptr;     // [parity 2][grid]
  // subgrid face fluxes (hipPPMLevelSetSubgridFluxes): per grid, per axis
  // the face records, their owners and the device staging
  struct Faces {
    int nsub = 0;
    std::vector<FaceRec> recs[3];
    std::vector<int> sub[3], side[3];
    long long total[3] = {0, 0, 0};
    int *head[3] = {nullptr, nullptr, nullptr};
    FaceRec *recs_dev[3] = {nullptr, nullptr, nullptr};
    double *out[3] = {nullptr, nullptr, nullptr};
  };
  std::vector<Faces> faces;
  // sibling copies (hipPPMLevelSetSiblings)
  CopyBox *dboxes = nullptr;
  GridGeom *dgeom = nullptr;
  long long nboxes = 0, max_box = 0;
};

namespace {

void drop_graphs(hipPPMLevel *L) {
  for (auto &kv : L->graphs) (void)hipGraphExecDestroy(kv.second);
  L->graphs.clear();
}

void free_faces(hipPPMLevel::Faces &F) {
  for (int d = 0; d < 3; d++) {
    if (F.head[d]) (void)hipFree(F.head[d]);
    if (F.recs_dev[d]) (void)hipFree(F.recs_dev[d]);
    if (F.out[d]) (void)hipFree(F.out[d]);
    F.head[d] = nullptr;
    F.recs_dev[d] = nullptr;
    F.out[d] = nullptr;
    F.total[d] = 0;
    F.recs[d].clear();
    F.sub[d].clear();
    F.side[d].clear();
  }
  F.nsub = 0;
}

// a grid's sweep parameters point at its face records (or at none)
void face_params(const hipPPMLevel *L, int g, int d, SweepParams &S) {
  const bool on = g < (int)L->faces.size() && L->faces[g].total[d] > 0;
  S.face_head = on ? L->faces[g].head[d] : nullptr;
  S.faces = on ? L->faces[g].recs_dev[d] : nullptr;
  S.face_out = on ? L->faces[g].out[d] : nullptr;
}

void free_level(hipPPMLevel *L) {
  drop_graphs(L);
  for (auto &F : L->faces) free_faces(F);
  void *ptrs[] = {L->A, L->B, L->pres, L->accel, L->tabs, L->cws, L->flat, L->qc, L->segs, L->dt,
                  L->dtmin, L->counters, L->fb_flag, L->fb_list, L->fb_state, L->grp_grid,
                  L->dgrp_off, L->dparams,
                  L->dflat, L->dpres, L->dwork, L->dpwork, L->learn, L->cws_dt, L->old,
                  L->dexp, L->dboxes, L->dgeom, L->dfball, L->queues};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (L->own_stream && L->stream) (void)hipStreamDestroy(L->stream);
}

// the level's switches must agree grid to grid
bool same_switches(const hipPPMParameter &a, const hipPPMParameter &b) {
  if (a.GridRank != b.GridRank || a.NumberOfColours != b.NumberOfColours) return false;
  if (a.DensNum != b.DensNum || a.TENum != b.TENum || a.Vel1Num != b.Vel1Num ||
      a.Vel2Num != b.Vel2Num || a.Vel3Num != b.Vel3Num || a.GENum != b.GENum)
    return false;
  for (int n = 0; n < a.NumberOfColours; n++)
    if (a.colnum[n] != b.colnum[n]) return false;
  return a.Gamma == b.Gamma && a.PPMDiffusionParameter == b.PPMDiffusionParameter &&
         a.PPMFlatteningParameter == b.PPMFlatteningParameter &&
         a.PPMSteepeningParameter == b.PPMSteepeningParameter && a.GravityOn == b.GravityOn &&
         a.DualEnergyFormalism == b.DualEnergyFormalism &&
         a.DualEnergyFormalismEta1 == b.DualEnergyFormalismEta1 &&
         a.DualEnergyFormalismEta2 == b.DualEnergyFormalismEta2 &&
         a.PressureFree == b.PressureFree && a.MinimumPressure == b.MinimumPressure &&
         a.RiemannSolver == b.RiemannSolver && a.RiemannSolverFallback == b.RiemannSolverFallback &&
         a.SmallRho == b.SmallRho && a.ConservativeReconstruction == b.ConservativeReconstruction &&
         a.PositiveReconstruction == b.PositiveReconstruction;
}

// sweep orientation: t1/t2 axes (t1 the contiguous transverse axis, or the
// row axis with diffusion), as the per-grid library does
void orient(int dim, bool diffusion, int &ax1, int &ax2) {
  const int rax = (dim + 1) % 3, oax = (dim + 2) % 3;
  if (diffusion) { ax1 = rax; ax2 = oax; return; }
  if (dim == 0) { ax1 = 1; ax2 = 2; }
  else if (dim == 1) { ax1 = 0; ax2 = 2; }
  else { ax1 = 0; ax2 = 1; }
}

// sweep widths (a_sweep * CellWidth when comoving, the product as
// Grid_SolveHydroEquations.C:386-408 takes it), their coefficient tables and
// uniform-axis constants, uploaded; `uni_changed` tells whether an axis
// changed between uniform and not (the launches depend on that)
int upload_geometry(hipPPMLevel *L, bool *uni_changed = nullptr) {
  std::vector<double> htabs, hcw(L->hcw.size());
  std::vector<int> uni(L->uni.size());
  std::vector<double> kt(L->kt.size(), 0.0);
  for (int g = 0; g < L->ngrids; g++)
    for (int d = 0; d < 3; d++) {
      const int N = L->GD[3 * g + d];
      const long long o = L->cw_off[3 * g + d];
      std::vector<double> dx(N), tab;
      for (int i = 0; i < N; i++)
        dx[i] = L->comoving && d < L->rank ? L->a_sweep * L->hcw[o + i] : L->hcw[o + i];
      build_tables(dx.data(), N, tab);
      uni[3 * g + d] = uniform_axis(dx.data(), N, tab, &kt[((size_t)3 * g + d) * NTAB]) ? 1 : 0;
      htabs.insert(htabs.end(), tab.begin(), tab.end());
      std::copy(dx.begin(), dx.end(), hcw.begin() + o);
    }
  if (uni_changed) *uni_changed = uni != L->uni;
  L->uni = uni;
  L->kt = kt;
  LCHK(hipMemcpyAsync(L->tabs, htabs.data(), sizeof(double) * htabs.size(), hipMemcpyHostToDevice,
                      L->stream));
  LCHK(hipMemcpyAsync(L->cws, hcw.data(), sizeof(double) * hcw.size(), hipMemcpyHostToDevice,
                      L->stream));
  LCHK(hipStreamSynchronize(L->stream));
  return ENZO_PPM_SUCCESS;
}

// pressure/Courant and expansion parameters of every grid, both parities
int upload_params(hipPPMLevel *L, double coef = 1.0, double Coefficient = 0.0) {
  const int G = L->ngrids;
  std::vector<PresParams> hp((size_t)2 * G);
  std::vector<ExpandParams> hx((size_t)2 * G);
  for (int par = 0; par < 2; par++)
    for (int g = 0; g < G; g++) {
      PresParams &Q = hp[(size_t)par * G + g];
      memset(&Q, 0, sizeof(Q));
      const hipPPMParameter &P = L->para[g];
      Q.rank = L->rank;
      Q.idual = (int)P.DualEnergyFormalism;
      for (int d = 0; d < 3; d++) {
        Q.GD[d] = L->GD[3 * g + d];
        Q.st[d] = (int)P.GridStartIndex[d];
        Q.en[d] = (int)P.GridEndIndex[d];
        Q.cw[d] = (L->comoving ? L->cws_dt : L->cws) + L->cw_off[3 * g + d];
      }
      Q.cells = L->cells[g];
      Q.A = (par == 0 ? L->A : L->B) + L->off[g];
      Q.pres = L->pres + L->poff[g];
      Q.gamma = P.Gamma;
      Q.tiny_number = P.MinimumPressure;
      Q.min_support = 0.0;
      Q.aye = L->comoving ? L->aye : 1.0;
      Q.want_dt = 1;
      Q.dtmin = L->dtmin;
      ExpandParams &X = hx[(size_t)par * G + g];
      memset(&X, 0, sizeof(X));
      X.cells = L->cells[g];
      X.rank = L->rank;
      X.idual = (int)P.DualEnergyFormalism;
      X.interp = L->old_valid && coef != 1.0 ? 1 : 0;
      X.gamma = P.Gamma;
      X.tiny_number = P.MinimumPressure;
      X.coef = coef;
      X.Coefficient = Coefficient;
      X.A = (par == 0 ? L->A : L->B) + L->off[g];
      X.old = L->old ? L->old + L->ooff[g] : nullptr;
    }
  LCHK(hipMemcpyAsync(L->dpres, hp.data(), sizeof(PresParams) * hp.size(), hipMemcpyHostToDevice,
                      L->stream));
  LCHK(hipMemcpyAsync(L->dexp, hx.data(), sizeof(ExpandParams) * hx.size(),
                      hipMemcpyHostToDevice, L->stream));
  LCHK(hipStreamSynchronize(L->stream));
  return ENZO_PPM_SUCCESS;
}

// SweepParams, FlatParams and work lists for the current shape choice
int build_plan(hipPPMLevel *L) {
  const hipPPMParameter &P0 = L->para0;
  const int G = L->ngrids;
  L->hparams.assign((size_t)2 * 3 * G, SweepParams());
  std::vector<FlatParams> hflat((size_t)2 * 3 * G);
  std::vector<Work> work;
  std::vector<unsigned> fball;
  long long ngroups = 0, nlaunch = 0;
  L->grp_off.assign(G, 0);
  for (int g = 0; g < G; g++) {  // flags per grid: the largest group count of any axis
    long long m = 1;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        if (a != b) m = std::max(m, (long long)((L->GD[3 * g + a] + 3) / 4) * L->GD[3 * g + b]);
    L->grp_off[g] = ngroups;
    ngroups += m;
  }
  if (ngroups != L->ngroups) return ENZO_PPM_ERR_ARGUMENT;
  for (int d = 0; d < 3; d++) {
    L->launches[d].clear();
    L->launch_slices[d].clear();
    L->prep_blocks[d] = 1;
  }
  for (int d = 0; d < L->rank; d++) {
    // parameters of every grid for axis d
    int max_slices = 1;
    for (int g = 0; g < G; g++) {
      const int *gd = &L->GD[3 * g];
      const int rax = (d + 1) % 3, oax = (d + 2) % 3;
      const hipPPMParameter &Pg = L->para[g];
      const int s = L->shape[3 * g + d];
      for (int par = 0; par < 2; par++) {
        SweepParams &S = L->hparams[((size_t)par * 3 + d) * G + g];
        memset(&S, 0, sizeof(S));
        S.dim = d;
        S.N = gd[d];
        S.i1 = (int)Pg.GridStartIndex[d] + 1;
        S.i2 = (int)Pg.GridEndIndex[d] + 1;
        S.rank = L->rank;
        for (int k = 0; k < 3; k++) S.GD[k] = gd[k];
        S.cells = L->cells[g];
        orient(d, L->diffusion, S.ax_t1, S.ax_t2);
        S.n_t1 = gd[S.ax_t1];
        S.n_t2 = gd[S.ax_t2];
        S.t2_0 = 0;
        set_strides(S);
        S.tp = kShapeTP[s];
        S.nblk_t1 = (S.n_t1 + S.tp - 1) / S.tp;
        const long long so = L->seg_off[((size_t)g * 3 + d) * kShapes + s];
        S.nseg = L->seg_n[((size_t)g * 3 + d) * kShapes + s];
        S.seg_o = L->segs + so;
        S.seg_w = L->segs + so + S.nseg;
        double *src = par == 0 ? L->A : L->B, *dst = par == 0 ? L->B : L->A;
        S.A = src + L->off[g];
        S.B = dst + L->off[g];
        S.ncol = L->ncol;
        S.learn = L->learned ? L->learn : nullptr;
        S.pres = L->pres + L->poff[g];
        S.gr = L->gravity ? L->accel + 3 * L->poff[g] + (long long)d * L->cells[g] : nullptr;
        S.flat = L->flat + L->flat_off[g];
        S.tab = L->tabs + L->tab_off[3 * g + d];
        S.tab_row = L->tabs + L->tab_off[3 * g + rax];
        S.tab_out = L->tabs + L->tab_off[3 * g + oax];
        S.qc = L->qc + L->qc_off[g];
        S.uni[0] = L->uni[3 * g + d];
        S.uni[1] = L->uni[3 * g + rax];
        S.uni[2] = L->uni[3 * g + oax];
        for (int t = 0; t < NTAB; t++) {
          S.kt[t] = L->kt[((size_t)3 * g + d) * NTAB + t];
          S.kt_row[t] = L->kt[((size_t)3 * g + rax) * NTAB + t];
          S.kt_out[t] = L->kt[((size_t)3 * g + oax) * NTAB + t];
        }
        set_dx_inverse(S);
        S.gamma = P0.Gamma;
        S.eta1 = P0.DualEnergyFormalismEta1;
        S.eta2 = P0.DualEnergyFormalismEta2;
        S.pmin = P0.MinimumPressure;
        S.dfloor = P0.SmallRho;
        S.dt = 0.0;
        S.dtp = L->dt;
        S.isteep = (int)P0.PPMSteepeningParameter;
        S.iflatten = (int)P0.PPMFlatteningParameter;
        S.idiff = (int)P0.PPMDiffusionParameter;
        S.idual = (int)P0.DualEnergyFormalism;
        S.ipresfree = (int)P0.PressureFree;
        S.gravity = (int)P0.GravityOn;
        S.ifallback = (int)P0.RiemannSolverFallback;
        S.nact_out = (int)(Pg.GridEndIndex[oax] - Pg.GridStartIndex[oax] + 1);
        face_params(L, g, d, S);
        S.counters = L->counters;
        S.fb_flag = L->fb_flag + L->grp_off[g];
        S.fb_list = L->fb_list;
        S.fb_count = L->fb_state;
        S.fb_done = L->fb_state + 1;
        S.fb_tag = (unsigned)L->grp_off[g];
        S.fb_ok = pencil_fits<kFbTS>(S) ? 1 : 0;
        S.hll_all = P0.RiemannSolver == ENZO_PPM_RIEMANN_HLL ? 1 : 0;
        S.hllc = P0.RiemannSolver == ENZO_PPM_RIEMANN_HLLC ? 1 : 0;
        FlatParams &F = hflat[((size_t)par * 3 + d) * G + g];
        memset(&F, 0, sizeof(F));
        F.dim = d;
        F.N = gd[d];
        F.i1 = S.i1;
        F.i2 = S.i2;
        for (int k = 0; k < 3; k++) F.GD[k] = gd[k];
        F.cells = L->cells[g];
        F.n_outer = gd[oax];
        F.iflatten = S.iflatten;
        F.gamma = S.gamma;
        F.A = S.A;
        F.pres = S.pres;
        F.flat = L->flat + L->flat_off[g];
        F.dx = L->cws + L->cw_off[3 * g + d];
        F.dt = 0.0;
        F.dtp = L->dt;
        F.qc = L->qc + L->qc_off[g];
        const long long items =
            F.N + (F.iflatten ? (long long)(F.i2 - F.i1 + 3) * F.n_outer : 0);
        L->prep_blocks[d] = (int)std::max<long long>(
            L->prep_blocks[d], std::min<long long>((items + 255) / 256, 1024));
      }
      if (L->diffusion) max_slices = std::max(max_slices, gd[(d + 2) % 3]);
    }
    // work lists: per slice (diffusion) or one, then per (shape, uni)
    const int nslices = L->diffusion ? max_slices : 1;
    L->fball_first[d].clear();
    for (int k = 0; k < nslices; k++) {
      L->fball_first[d].push_back((long long)fball.size());
      if (L->whole)
        for (int g = 0; g < G; g++) {
          const SweepParams &S = L->hparams[((size_t)0 * 3 + d) * G + g];
          if (L->diffusion) {
            if (k >= S.n_t2) continue;
            for (int r = 0; r < S.nblk_t1; r++)
              fball.push_back((unsigned)(L->grp_off[g] + r + (long long)S.nblk_t1 * k));
          } else {
            for (long long r = 0; r < (long long)S.nblk_t1 * S.n_t2; r++)
              fball.push_back((unsigned)(L->grp_off[g] + r));
          }
        }
      L->launch_slices[d].push_back((long long)L->launches[d].size());
      for (int s = 0; s < kShapes; s++)
        for (int u = 1; u >= 0; u--) {
          const long long first = (long long)work.size();
          for (int g = 0; g < G; g++) {
            if (L->shape[3 * g + d] != s || L->uni[3 * g + d] != u) continue;
            const SweepParams &S = L->hparams[((size_t)0 * 3 + d) * G + g];
            if (L->diffusion) {
              if (k >= S.n_t2) continue;
              for (int r = 0; r < S.nblk_t1; r++)
                for (int seg = 0; seg < S.nseg; seg++)
                  work.push_back(Work{g, seg + S.nseg * (r + S.nblk_t1 * k)});
            } else {
              const long long nt = (long long)S.nseg * S.nblk_t1 * S.n_t2;
              for (long long t = 0; t < nt; t++) work.push_back(Work{g, (int)t});
            }
          }
          const long long cnt = (long long)work.size() - first;
          if (cnt > 0) L->launches[d].push_back(hipPPMLevel::Launch{s, u, first, cnt, nlaunch++});
        }
    }
    L->launch_slices[d].push_back((long long)L->launches[d].size());
    L->fball_first[d].push_back((long long)fball.size());
  }
  for (int d = 0; d < 3; d++)
    for (int ts : kFbWidths) {  // the narrowest one-pencil width holding every grid's pencils
      L->fb_ts[d] = ts;
      bool all = true;
      for (int g = 0; g < G && all; g++)
        all = pencil_fits_width(L->hparams[((size_t)0 * 3 + d) * G + g], ts);
      if (all) break;
    }
  // host check of every item against its grid's tiles
  for (const Work &w : work) {
    if (w.grid < 0 || w.grid >= G) return ENZO_PPM_ERR_ARGUMENT;
  }
  for (int d = 0; d < L->rank; d++)
    for (const auto &ln : L->launches[d])
      for (long long i = ln.first; i < ln.first + ln.count; i++) {
        const Work &w = work[i];
        const SweepParams &S = L->hparams[((size_t)0 * 3 + d) * G + w.grid];
        if (w.tile < 0 || (long long)w.tile >= (long long)S.nseg * S.nblk_t1 * S.n_t2)
          return ENZO_PPM_ERR_ARGUMENT;
      }
  if (L->dparams) (void)hipFree(L->dparams);
  if (L->dflat) (void)hipFree(L->dflat);
  if (L->dwork) (void)hipFree(L->dwork);
  L->dparams = nullptr;
  L->dflat = nullptr;
  L->dwork = nullptr;
  if (dev_alloc(&L->dparams, L->hparams.size()) || dev_alloc(&L->dflat, hflat.size()) ||
      dev_alloc(&L->dwork, work.size()))
    return ENZO_PPM_ERR_HIP;
  LCHK(hipMemcpy(L->dparams, L->hparams.data(), sizeof(SweepParams) * L->hparams.size(),
                 hipMemcpyHostToDevice));
  LCHK(hipMemcpy(L->dflat, hflat.data(), sizeof(FlatParams) * hflat.size(), hipMemcpyHostToDevice));
  if (!work.empty())
    LCHK(hipMemcpy(L->dwork, work.data(), sizeof(Work) * work.size(), hipMemcpyHostToDevice));
  // whole-pencil group lists: every entry a group of the level, each list no
  // longer than the fallback list
  for (int d = 0; d < L->rank; d++)
    for (size_t k = 0; k + 1 < L->fball_first[d].size(); k++)
      if (L->fball_first[d][k + 1] - L->fball_first[d][k] > L->ngroups) return ENZO_PPM_ERR_ARGUMENT;
  for (unsigned e : fball)
    if ((long long)e >= L->ngroups) return ENZO_PPM_ERR_ARGUMENT;
  if (L->dfball) (void)hipFree(L->dfball);
  L->dfball = nullptr;
  if (dev_alloc(&L->dfball, fball.size())) return ENZO_PPM_ERR_HIP;
  // the persistent launches' counters start at 0 (and their last workgroups
  // leave them at 0)
  if (L->queues) (void)hipFree(L->queues);
  L->queues = nullptr;
  if (dev_alloc(&L->queues, (size_t)std::max(1LL, nlaunch) * 2 * kXCD)) return ENZO_PPM_ERR_HIP;
  LCHK(hipMemset(L->queues, 0, sizeof(unsigned) * (size_t)std::max(1LL, nlaunch) * 2 * kXCD));
  if (!fball.empty())
    LCHK(hipMemcpy(L->dfball, fball.data(), sizeof(unsigned) * fball.size(), hipMemcpyHostToDevice));
  drop_graphs(L);
  return ENZO_PPM_SUCCESS;
}

// the level's fallback launch in one-pencil workgroups of width ts,
// blocks1024 * kFbTS work-items in all
// and at most `ngroups` workgroups (the most entries the list can hold)
hipError_t launch_level_fallback(hipPPMLevel *L, const SweepParams *P, int ts, int blocks1024,
                                 long long ngroups) {
  const unsigned nblk =
      (unsigned)std::max(1LL, std::min(ngroups, (long long)fb_blocks(ts, blocks1024 * kFbTS)));
  return with_fb_width(ts, [&](auto W) {
    constexpr int TS = decltype(W)::value;
    if (L->learned)
      hipLaunchKernelGGL((level_fallback_kernel<1, TS>), dim3(nblk), dim3(TS), 0, L->stream, P,
                         L->fb_list, L->grp_grid, L->dgrp_off, L->fb_state, L->fb_state + 1);
    else
      hipLaunchKernelGGL((level_fallback_kernel<0, TS>), dim3(nblk), dim3(TS), 0, L->stream, P,
                         L->fb_list, L->grp_grid, L->dgrp_off, L->fb_state, L->fb_state + 1);
  });
}

int launch_level_sweep(hipPPMLevel *L, int d) {
  const int G = L->ngrids;
  const SweepParams *P = L->dparams + ((size_t)L->parity * 3 + d) * G;
  hipLaunchKernelGGL(level_prep_kernel, dim3((unsigned)L->prep_blocks[d], (unsigned)G), dim3(256),
                     0, L->stream,
                     L->dflat + ((size_t)L->parity * 3 + d) * G);
  LCHK(hipGetLastError());
  const auto &sl = L->launch_slices[d];
  for (size_t k = 0; k + 1 < sl.size(); k++) {
    if (L->whole) {
      // HLL / HLLC: every group of the slice to the one-pencil workgroups
      const long long f0 = L->fball_first[d][k], n = L->fball_first[d][k + 1] - f0;
      if (n == 0) continue;
      hipLaunchKernelGGL(level_fb_fill_kernel, dim3((unsigned)std::min<long long>((n + 255) / 256, 1024)),
                         dim3(256), 0, L->stream, L->dfball + f0, (unsigned)n, L->fb_list, L->fb_state);
      LCHK(hipGetLastError());
      LCHK(launch_level_fallback(L, P, L->fb_ts[d], kFbBlocksWhole, n));
      continue;
    }
    for (long long li = sl[k]; li < sl[k + 1]; li++) {
      const auto &ln = L->launches[d][li];
      const Work *w = L->dwork + ln.first;
      const dim3 grid((unsigned)ln.count), blk(256);
      const dim3 pgrid((unsigned)std::min<long long>(ln.count, L->persistent_blocks));
      unsigned *queue = L->queues + (size_t)2 * kXCD * ln.queue;
      const int nit = (int)ln.count;
#define LEVEL_LAUNCH1(TP, TS, U, LRN)                                                         \
  if (L->persistent)                                                                        \
    hipLaunchKernelGGL((level_sweep_persistent_kernel<TP, TS, U, LRN>), pgrid, blk, 0,       \
                       L->stream, P, w, nit, queue);                                         \
  else                                                                                      \
    hipLaunchKernelGGL((level_sweep_kernel<TP, TS, U, LRN>), grid, blk, 0, L->stream, P, w);
#define LEVEL_LAUNCH(S, TP, TS)                                                             \
  if (ln.shape == S) {                                                                    \
    if (L->learned) {                                                                     \
      if (ln.uni) { LEVEL_LAUNCH1(TP, TS, 1, 1) } else { LEVEL_LAUNCH1(TP, TS, 0, 1) }   \
    } else {                                                                              \
      if (ln.uni) { LEVEL_LAUNCH1(TP, TS, 1, 0) } else { LEVEL_LAUNCH1(TP, TS, 0, 0) }   \
    }                                                                                     \
  }
      LEVEL_LAUNCH(0, 16, 16)
      LEVEL_LAUNCH(1, 8, 32)
      LEVEL_LAUNCH(2, 4, 64)
#undef LEVEL_LAUNCH
#undef LEVEL_LAUNCH1
      LCHK(hipGetLastError());
    }
    LCHK(launch_level_fallback(L, P, L->fb_ts[d], kFbBlocks, L->ngroups));
  }
  L->parity ^= 1;
  return ENZO_PPM_SUCCESS;
}

// one level step: pressure (+ Courant dt) of every grid -- unless
// `pressure_current` (the pressure array already holds this state's
// pressure and dt is given) -- then the sweeps in the rotated order of
// CycleNumber
int enqueue_step(hipPPMLevel *L, int cycle, int want_dt, double cfl, bool pressure_current) {
  if (L->comoving) {
    hipLaunchKernelGGL(level_save_old_kernel, dim3((unsigned)L->npwork), dim3(256), 0, L->stream,
                       L->dexp + (size_t)L->parity * L->ngrids, L->dpwork);
    LCHK(hipGetLastError());
  }
  if (want_dt)
    hipLaunchKernelGGL(level_dt_init_kernel, dim3(1), dim3(64), 0, L->stream, L->dtmin);
  if (want_dt || !pressure_current)
    hipLaunchKernelGGL(level_pressure_kernel, dim3((unsigned)L->npwork), dim3(256), 0, L->stream,
                       L->dpres + (size_t)L->parity * L->ngrids, L->dpwork);
  LCHK(hipGetLastError());
  if (want_dt)
    hipLaunchKernelGGL(level_dt_final_kernel, dim3(1), dim3(64), 0, L->stream, L->dtmin, cfl, L->dt);
  const int ixyz = cycle % L->rank;
  for (int n = ixyz; n < ixyz + L->rank; n++) {
    int st = launch_level_sweep(L, n % L->rank);
    if (st != ENZO_PPM_SUCCESS) return st;
  }
  return ENZO_PPM_SUCCESS;
}

}  // namespace

extern "C" {

int hipPPMLevelCreate(hipPPMLevel **Level, enzo_int NumberOfGrids, const hipPPMParameter *Para,
                      void *stream) {
  if (!Level || !Para || NumberOfGrids < 1 || NumberOfGrids > 65535) return ENZO_PPM_ERR_ARGUMENT;
  *Level = nullptr;
  for (enzo_int g = 0; g < NumberOfGrids; g++) {
    const hipPPMParameter &P = Para[g];
    if (P.GridRank < 1 || P.GridRank > 3 || P.NumberOfColours < 0 ||
        P.NumberOfColours > ENZO_PPM_MAX_COLOR || !same_switches(P, Para[0]))
      return ENZO_PPM_ERR_ARGUMENT;
    for (int d = 0; d < 3; d++) {
      if (P.GridDimension[d] < 1 || P.GridDimension[d] > ENZO_PPM_MAX_ANY_SINGLE_DIRECTION)
        return ENZO_PPM_ERR_ARGUMENT;
      // every swept axis has more than one active cell: the level swaps its
      // buffers for all grids at once, so no grid may skip a sweep
      if (d < P.GridRank && (P.GridStartIndex[d] < 3 || P.GridEndIndex[d] + 3 >= P.GridDimension[d] ||
                             P.GridEndIndex[d] <= P.GridStartIndex[d]))
        return ENZO_PPM_ERR_ARGUMENT;
    }
    if (P.ConservativeReconstruction != 0 || P.PositiveReconstruction != 0 ||
        (P.RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK && P.RiemannSolver != ENZO_PPM_RIEMANN_HLL &&
         P.RiemannSolver != ENZO_PPM_RIEMANN_HLLC) ||
        P.PPMFlatteningParameter < 0 || P.PPMFlatteningParameter > 3 ||
        (P.PPMDiffusionParameter != 0 && P.PPMDiffusionParameter != 1))
      return ENZO_PPM_ERR_NOT_SUPPORTED;
    // HLL / HLLC run in one-pencil workgroups: every swept pencil must fit
    if (P.RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK)
      for (int d = 0; d < P.GridRank; d++)
        if (P.GridDimension[d] > kFbTS + 1) return ENZO_PPM_ERR_NOT_SUPPORTED;
  }
  hipPPMLevel *L = new (std::nothrow) hipPPMLevel();
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  for (double &t : L->tune_times) t = std::nan("");
  const int G = (int)NumberOfGrids;
  L->ngrids = G;
  L->para.assign(Para, Para + G);
  L->para0 = Para[0];
  L->rank = (int)Para[0].GridRank;
  L->ncol = (int)Para[0].NumberOfColours;
  L->nslot = S_COL0 + L->ncol;
  L->gravity = Para[0].GravityOn != 0;
  L->diffusion = Para[0].PPMDiffusionParameter != 0;
  L->whole = Para[0].RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK ? 1 : 0;
  L->faces.resize(G);
  L->GD.resize(3 * G);
  L->cells.resize(G);
  L->off.resize(G);
  L->poff.resize(G);
  L->tab_off.resize(3 * G);
  L->cw_off.resize(3 * G);
  L->flat_off.resize(G);
  L->qc_off.resize(G);
  L->uni.resize(3 * G);
  L->kt.assign((size_t)3 * G * NTAB, 0.0);
  L->shape.assign(3 * G, 1);
  L->seg_off.assign((size_t)G * 3 * kShapes, 0);
  L->seg_n.assign((size_t)G * 3 * kShapes, 0);
  long long tot = 0, ptot = 0, ttot = 0, wtot = 0, ftot = 0, qtot = 0;
  for (int g = 0; g < G; g++) {
    const hipPPMParameter &P = Para[g];
    long long c = 1;
    for (int d = 0; d < 3; d++) {
      L->GD[3 * g + d] = (int)P.GridDimension[d];
      c *= P.GridDimension[d];
    }
    L->cells[g] = c;
    L->off[g] = tot;
    L->poff[g] = ptot;
    tot += c * L->nslot;
    ptot += c;
    long long fcap = 0;
    int nmax = 1;
    for (int d = 0; d < 3; d++) {
      const int N = L->GD[3 * g + d];
      L->tab_off[3 * g + d] = ttot;
      ttot += (long long)NTAB * N;
      L->cw_off[3 * g + d] = wtot;
      for (int i = 0; i < N; i++)
        L->hcw.push_back(P.CellWidth[d] ? P.CellWidth[d][i] : (d < P.GridRank ? P.dx[d] : 1.0));
      wtot += N;
      fcap = std::max(fcap, (long long)N * L->GD[3 * g + (d + 2) % 3]);
      nmax = std::max(nmax, N);
    }
    L->flat_off[g] = ftot;
    ftot += fcap;
    L->qc_off[g] = qtot;
    qtot += nmax;
  }
  // segments of every (grid, axis, shape); a shape that cannot tile an axis
  // is marked with nseg 0 and never chosen
  for (int g = 0; g < G; g++)
    for (int d = 0; d < L->rank; d++) {
      const int N = L->GD[3 * g + d];
      const int i1 = (int)Para[g].GridStartIndex[d] + 1, i2 = (int)Para[g].GridEndIndex[d] + 1;
      for (int s = 0; s < kShapes; s++) {
        std::vector<int> o, w;
        const size_t k = ((size_t)g * 3 + d) * kShapes + s;
        L->seg_off[k] = (long long)L->hsegs.size();
        if (!segments_shape(s, N, i1, i2, o, w)) continue;
        L->seg_n[k] = (int)o.size();
        L->hsegs.insert(L->hsegs.end(), o.begin(), o.end());
        L->hsegs.insert(L->hsegs.end(), w.begin(), w.end());
      }
      if (L->seg_n[((size_t)g * 3 + d) * kShapes + 1] == 0) {
        delete L;
        return ENZO_PPM_ERR_ARGUMENT;
      }
    }
  long long ngroups = 0;
  std::vector<int> hgrp_grid;
  std::vector<unsigned> hgrp_off(G);
  for (int g = 0; g < G; g++) {
    long long m = 1;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        if (a != b) m = std::max(m, (long long)((L->GD[3 * g + a] + 3) / 4) * L->GD[3 * g + b]);
    hgrp_off[g] = (unsigned)ngroups;
    hgrp_grid.insert(hgrp_grid.end(), (size_t)m, g);
    ngroups += m;
  }
  if (ngroups >= (1ll << 32)) {
    delete L;
    return ENZO_PPM_ERR_ARGUMENT;
  }
  L->ngroups = ngroups;
  // pressure chunks
  std::vector<Work> pw;
  for (int g = 0; g < G; g++)
    for (long long s = 0; s < L->cells[g]; s += kPresChunk) pw.push_back(Work{g, (int)s});
  L->npwork = (long long)pw.size();
  int st = ENZO_PPM_SUCCESS;
  auto fail = [&](int code) {
    free_level(L);
    delete L;
    return code;
  };
  if (dev_alloc(&L->A, tot) || dev_alloc(&L->B, tot) || dev_alloc(&L->pres, ptot) ||
      (L->gravity && dev_alloc(&L->accel, 3 * ptot)) || dev_alloc(&L->tabs, ttot) ||
      dev_alloc(&L->cws, wtot) || dev_alloc(&L->flat, ftot) || dev_alloc(&L->qc, qtot) ||
      dev_alloc(&L->segs, L->hsegs.size()) || dev_alloc(&L->dt, 1) || dev_alloc(&L->dtmin, 1) ||
      dev_alloc(&L->counters, 4) || dev_alloc(&L->fb_flag, ngroups) ||
      dev_alloc(&L->fb_list, ngroups) || dev_alloc(&L->fb_state, 2) ||
      dev_alloc(&L->grp_grid, ngroups) || dev_alloc(&L->dgrp_off, G) ||
      dev_alloc(&L->dpres, (size_t)2 * G) || dev_alloc(&L->dpwork, pw.size()))
    return fail(ENZO_PPM_ERR_HIP);
  if (stream) {
    L->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&L->stream) != hipSuccess) return fail(ENZO_PPM_ERR_HIP);
    L->own_stream = true;
  }
  {  // persistent launches: 4 workgroups per CU (the sweep tiles' occupancy)
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        ncu > 0)
      L->persistent_blocks = 4 * ncu;
  }
  if (upload_geometry(L) != ENZO_PPM_SUCCESS) return fail(ENZO_PPM_ERR_HIP);
  if ((!L->hsegs.empty() && hipMemcpy(L->segs, L->hsegs.data(), sizeof(int) * L->hsegs.size(),
                                      hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(L->dpwork, pw.data(), sizeof(Work) * pw.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(L->fb_flag, 0, sizeof(int) * ngroups) != hipSuccess ||
      hipMemcpy(L->grp_grid, hgrp_grid.data(), sizeof(int) * ngroups, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(L->dgrp_off, hgrp_off.data(), sizeof(unsigned) * G, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(L->fb_state, 0, 2 * sizeof(unsigned)) != hipSuccess ||
      hipMemset(L->counters, 0, 4 * sizeof(unsigned long long)) != hipSuccess)
    return fail(ENZO_PPM_ERR_HIP);
  if (L->gravity && hipMemset(L->accel, 0, sizeof(double) * 3 * ptot) != hipSuccess)
    return fail(ENZO_PPM_ERR_HIP);
  if (dev_alloc(&L->dexp, (size_t)2 * G) || upload_params(L) != ENZO_PPM_SUCCESS)
    return fail(ENZO_PPM_ERR_HIP);
  st = build_plan(L);
  if (st != ENZO_PPM_SUCCESS) return fail(st);
  *Level = L;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelSetBaryon(hipPPMLevel *L, enzo_int grid, enzo_float *BaryonField[],
                         enzo_float *AccelerationField[]) {
  if (!L || !BaryonField || grid < 0 || grid >= L->ngrids) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  const hipPPMParameter &P = L->para[grid];
  long long fnum[S_COL0 + ENZO_PPM_MAX_COLOR];
  for (int s = 0; s < L->nslot; s++) fnum[s] = -1;
  fnum[S_D] = P.DensNum;
  fnum[S_E] = P.TENum;
  fnum[S_VX] = P.Vel1Num;
  if (P.GridRank > 1) fnum[S_VY] = P.Vel2Num;
  if (P.GridRank > 2) fnum[S_VZ] = P.Vel3Num;
  if (P.DualEnergyFormalism) fnum[S_GE] = P.GENum;
  for (int n = 0; n < L->ncol; n++) fnum[S_COL0 + n] = P.colnum[n];
  double *state = L->parity == 0 ? L->A : L->B;
  const size_t fb = sizeof(double) * (size_t)L->cells[grid];
  for (int s = 0; s < L->nslot; s++) {
    double *dst = state + L->off[grid] + (long long)s * L->cells[grid];
    if (fnum[s] >= 0 && fnum[s] < ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS && BaryonField[fnum[s]])
      LCHK(hipMemcpyAsync(dst, BaryonField[fnum[s]], fb, hipMemcpyHostToDevice, L->stream));
    else
      LCHK(hipMemsetAsync(dst, 0, fb, L->stream));
  }
  if (L->gravity && AccelerationField)
    for (int d = 0; d < L->rank; d++)
      if (AccelerationField[d])
        LCHK(hipMemcpyAsync(L->accel + 3 * L->poff[grid] + (long long)d * L->cells[grid],
                            AccelerationField[d], fb, hipMemcpyHostToDevice, L->stream));
  LCHK(hipStreamSynchronize(L->stream));
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelGetBaryon(hipPPMLevel *L, enzo_int grid, enzo_float *BaryonField[]) {
  if (!L || !BaryonField || grid < 0 || grid >= L->ngrids) return ENZO_PPM_ERR_ARGUMENT;
  const hipPPMParameter &P = L->para[grid];
  long long fnum[S_COL0 + ENZO_PPM_MAX_COLOR];
  for (int s = 0; s < L->nslot; s++) fnum[s] = -1;
  fnum[S_D] = P.DensNum;
  fnum[S_E] = P.TENum;
  fnum[S_VX] = P.Vel1Num;
  if (P.GridRank > 1) fnum[S_VY] = P.Vel2Num;
  if (P.GridRank > 2) fnum[S_VZ] = P.Vel3Num;
  if (P.DualEnergyFormalism) fnum[S_GE] = P.GENum;
  for (int n = 0; n < L->ncol; n++) fnum[S_COL0 + n] = P.colnum[n];
  const double *state = L->parity == 0 ? L->A : L->B;
  const size_t fb = sizeof(double) * (size_t)L->cells[grid];
  for (int s = 0; s < L->nslot; s++)
    if (fnum[s] >= 0 && fnum[s] < ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS && BaryonField[fnum[s]])
      LCHK(hipMemcpyAsync(BaryonField[fnum[s]], state + L->off[grid] + (long long)s * L->cells[grid],
                          fb, hipMemcpyDeviceToHost, L->stream));
  LCHK(hipStreamSynchronize(L->stream));
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelStep(hipPPMLevel *L, enzo_int CycleNumber, enzo_float dtFixed, enzo_float cfl) {
  if (!L || CycleNumber < 0) return ENZO_PPM_ERR_ARGUMENT;
  const int want_dt = dtFixed > 0.0 ? 0 : 1;
  const bool pres_done = L->pres_current && !want_dt;
  L->pres_current = false;
  if (!want_dt) LCHK(hipMemcpyAsync(L->dt, &dtFixed, sizeof(double), hipMemcpyHostToDevice, L->stream));
  if (!L->use_graph) {
    L->old_valid = L->comoving;
    return enqueue_step(L, (int)(CycleNumber % L->rank), want_dt, cfl, pres_done);
  }
  // the launches of a step depend on the rotation, the buffer parity, the
  // dt mode and whether the pressure is already there
  const int rot = (int)(CycleNumber % L->rank);
  const int key = ((rot * 2 + L->parity) * 2 + want_dt) * 2 + (pres_done ? 1 : 0);
  if (want_dt && cfl != L->graph_cfl) {  // cfl is a kernel argument of those graphs
    for (auto g = L->graphs.begin(); g != L->graphs.end();) {
      if ((g->first / 2) % 2) {
        (void)hipGraphExecDestroy(g->second);
        g = L->graphs.erase(g);
      } else {
        ++g;
      }
    }
    L->graph_cfl = cfl;
  }
  auto it = L->graphs.find(key);
  if (it == L->graphs.end()) {
    const int parity0 = L->parity;
    hipGraph_t graph;
    LCHK(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue_step(L, rot, want_dt, cfl, pres_done);
    hipError_t ec = hipStreamEndCapture(L->stream, &graph);
    if (st != ENZO_PPM_SUCCESS) return st;
    LCHK(ec);
    hipGraphExec_t exec;
    LCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    it = L->graphs.emplace(key, exec).first;
    L->parity = parity0;  // capture ran no work
  }
  LCHK(hipGraphLaunch(it->second, L->stream));
  L->parity ^= (L->rank & 1);  // one swap per sweep
  L->old_valid = L->comoving;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelSetSubgridFluxes(hipPPMLevel *L, enzo_int grid, enzo_fluxes *SubgridFluxes[],
                                enzo_int NumberOfSubgrids, const enzo_long_int GridGlobalStart[]) {
  if (!L || grid < 0 || grid >= L->ngrids || NumberOfSubgrids < 0) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfSubgrids > 0 && (!SubgridFluxes || !GridGlobalStart)) return ENZO_PPM_ERR_ARGUMENT;
  const int g = (int)grid, G = L->ngrids;
  LCHK(hipStreamSynchronize(L->stream));  // no step may still use the old records
  hipPPMLevel::Faces &F = L->faces[g];
  free_faces(F);
  int st = ENZO_PPM_SUCCESS;
  for (int d = 0; d < L->rank && st == ENZO_PPM_SUCCESS; d++) {
    std::vector<int> head;
    long long total = 0;
    st = build_face_records(&L->GD[3 * g], &L->para[g], SubgridFluxes, (int)NumberOfSubgrids,
                            GridGlobalStart, L->nslot, d, F.recs[d], F.sub[d], F.side[d], head,
                            total);
    if (st != ENZO_PPM_SUCCESS || total == 0) continue;
    if (dev_alloc(&F.head[d], head.size()) || dev_alloc(&F.recs_dev[d], F.recs[d].size()) ||
        dev_alloc(&F.out[d], (size_t)total) ||
        hipMemcpy(F.head[d], head.data(), sizeof(int) * head.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(F.recs_dev[d], F.recs[d].data(), sizeof(FaceRec) * F.recs[d].size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(F.out[d], 0, sizeof(double) * total) != hipSuccess) {
      st = ENZO_PPM_ERR_HIP;
      continue;
    }
    F.total[d] = total;
  }
  if (st != ENZO_PPM_SUCCESS) free_faces(F);
  else F.nsub = (int)NumberOfSubgrids;
  // the grid's sweep parameters of both buffer parities; the kernels read
  // them from device memory at every launch, so captured steps see them too
  for (int par = 0; par < 2; par++)
    for (int d = 0; d < 3; d++) {
      const size_t k = ((size_t)par * 3 + d) * G + g;
      face_params(L, g, d, L->hparams[k]);
      LCHK(hipMemcpy(L->dparams + k, &L->hparams[k], sizeof(SweepParams), hipMemcpyHostToDevice));
    }
  return st;
}

int hipPPMLevelSaveSubgridFluxes(hipPPMLevel *L, enzo_int grid, enzo_fluxes *SubgridFluxes[],
                                 enzo_int NumberOfSubgrids) {
  if (!L || grid < 0 || grid >= L->ngrids) return ENZO_PPM_ERR_ARGUMENT;
  const hipPPMLevel::Faces &F = L->faces[grid];
  if (NumberOfSubgrids != F.nsub || (NumberOfSubgrids > 0 && !SubgridFluxes))
    return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  const hipPPMParameter &P = L->para[grid];
  const long long nact[3] = {P.GridEndIndex[0] - P.GridStartIndex[0] + 1,
                             P.GridEndIndex[1] - P.GridStartIndex[1] + 1,
                             P.GridEndIndex[2] - P.GridStartIndex[2] + 1};
  std::vector<long long> f;
  slot_fields(&P, f);
  for (int d = 0; d < L->rank; d++) {
    if (F.total[d] == 0) continue;
    std::vector<double> h((size_t)F.total[d]);
    LCHK(hipMemcpy(h.data(), F.out[d], sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    scatter_face_fluxes(h, F.recs[d], F.sub[d], F.side[d], f, nact, d, SubgridFluxes);
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelCorrectForRefinedFluxes(hipPPMLevel *L, enzo_int grid,
                                       const enzo_fluxes *InitialFluxes,
                                       const enzo_fluxes *RefinedFluxes,
                                       enzo_fluxes *BoundaryFluxesThisTimeStep,
                                       enzo_int SUBlingGrid, const enzo_int *GlobalDim,
                                       const enzo_int *PeriodicFace,
                                       enzo_int CorrectParentBoundaryFlux,
                                       const enzo_int *ColourKind) {
  if (!L || grid < 0 || grid >= L->ngrids) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  LCHK(hipStreamSynchronize(L->stream));  // the grid's state after every queued step
  double *state = (L->parity == 0 ? L->A : L->B) + L->off[grid];
  return correct_refined_fluxes(state, L->cells[grid], &L->GD[3 * grid], L->stream,
                                &L->para[grid], InitialFluxes, RefinedFluxes,
                                BoundaryFluxesThisTimeStep, SUBlingGrid, GlobalDim, PeriodicFace,
                                CorrectParentBoundaryFlux, ColourKind);
}

int hipPPMLevelStatus(hipPPMLevel *L, enzo_float *dt_used, enzo_long_int *fallback_cells) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  unsigned long long c[4];
  LCHK(hipMemcpy(c, L->counters, sizeof(c), hipMemcpyDeviceToHost));
  if (dt_used) LCHK(hipMemcpy(dt_used, L->dt, sizeof(double), hipMemcpyDeviceToHost));
  if (fallback_cells) *fallback_cells = (enzo_long_int)c[0];
  hipLaunchKernelGGL(level_zero_kernel, dim3(1), dim3(64), 0, L->stream, L->counters, 4);
  LCHK(hipGetLastError());
  LCHK(hipStreamSynchronize(L->stream));
  if (c[3] != 0) return ENZO_PPM_ERR_NOT_SUPPORTED;
  if (c[2] != 0) return ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelUseGraph(hipPPMLevel *L, enzo_int enable) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  L->use_graph = enable != 0;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelPersistent(hipPPMLevel *L, enzo_int enable) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  if ((enable == 0 || enable == 1) && L->persistent != (enable == 1)) {
    L->persistent = enable == 1;
    drop_graphs(L);  // the captured steps launch the other kernels
  }
  return L->persistent ? 1 : 0;
}

int hipPPMLevelSetShape(hipPPMLevel *L, enzo_int grid, enzo_int dim, enzo_int shape) {
  if (!L || grid < 0 || grid >= L->ngrids || dim < 0 || dim >= L->rank || shape < 0 ||
      shape >= kShapes)
    return ENZO_PPM_ERR_ARGUMENT;
  if (L->seg_n[((size_t)grid * 3 + dim) * kShapes + shape] == 0) return ENZO_PPM_ERR_ARGUMENT;
  L->shape[3 * grid + dim] = (int)shape;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelCommitShapes(hipPPMLevel *L) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  return build_plan(L);
}

int hipPPMLevelTune(hipPPMLevel *L, enzo_int CycleNumber, enzo_float dt, enzo_int repeats,
                    enzo_int *shape_by_length) {
  // Autotuner: pencils are classed by length (along the sweep axis, ghosts
  // included) into [0,20], [21,40], [41,80], [81,...]; for each class in turn
  // every tile shape that can tile all of the class's pencils is timed on
  // full level steps (graph replay, `repeats` of them) with the other classes
  // at their current choice, and the fastest kept.  The level's state is
  // saved before and restored after.  shape_by_length (4 entries) receives
  // the choice per class.
  if (!L || dt <= 0.0 || repeats < 1) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes (steps, then the saved state back)
  const int bounds[4] = {20, 40, 80, 1 << 30};
  auto cls = [&](int N) { int c = 0; while (N > bounds[c]) c++; return c; };
  long long tot = 0;
  for (int g = 0; g < L->ngrids; g++) tot += L->cells[g] * L->nslot;
  double *save = nullptr;
  if (dev_alloc(&save, tot)) return ENZO_PPM_ERR_HIP;
  LCHK(hipStreamSynchronize(L->stream));
  const int parity0 = L->parity;
  LCHK(hipMemcpy(save, parity0 == 0 ? L->A : L->B, sizeof(double) * tot, hipMemcpyDeviceToDevice));
  int choice[4];
  for (int c = 0; c < 4; c++) choice[c] = 1;
  for (double &t : L->tune_times) t = std::nan("");
  auto apply = [&]() {
    for (int g = 0; g < L->ngrids; g++)
      for (int d = 0; d < L->rank; d++) {
        int s = choice[cls(L->GD[3 * g + d])];
        if (L->seg_n[((size_t)g * 3 + d) * kShapes + s] == 0) s = 1;
        L->shape[3 * g + d] = s;
      }
    return build_plan(L);
  };
  int st = ENZO_PPM_SUCCESS;
  for (int c = 0; c < 4 && st == ENZO_PPM_SUCCESS; c++) {
    bool present = false;
    for (int g = 0; g < L->ngrids && !present; g++)
      for (int d = 0; d < L->rank; d++) present |= cls(L->GD[3 * g + d]) == c;
    if (!present) continue;
    double best = 1e300;
    int best_s = choice[c];
    for (int s = 0; s < kShapes; s++) {
      choice[c] = s;
      if ((st = apply()) != ENZO_PPM_SUCCESS) break;
      // warm-up (captures the graph), then timed repeats
      if ((st = hipPPMLevelStep(L, CycleNumber, dt, 0.0)) != ENZO_PPM_SUCCESS) break;
      if (hipStreamSynchronize(L->stream) != hipSuccess) { st = ENZO_PPM_ERR_HIP; break; }
      auto t0 = std::chrono::steady_clock::now();
      for (int r = 0; r < repeats && st == ENZO_PPM_SUCCESS; r++)
        st = hipPPMLevelStep(L, CycleNumber, dt, 0.0);
      if (hipStreamSynchronize(L->stream) != hipSuccess) st = ENZO_PPM_ERR_HIP;
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (st == ENZO_PPM_SUCCESS) L->tune_times[c * kShapes + s] = t / (double)repeats;
      if (st == ENZO_PPM_SUCCESS && t < best) {
        best = t;
        best_s = s;
      }
    }
    choice[c] = best_s;
  }
  if (st == ENZO_PPM_SUCCESS) st = apply();
  // then one workgroup per work item against the persistent launches
  if (st == ENZO_PPM_SUCCESS) {
    double tp[2] = {1e300, 1e300};
    for (int pers = 0; pers < 2 && st == ENZO_PPM_SUCCESS; pers++) {
      hipPPMLevelPersistent(L, pers);
      if ((st = hipPPMLevelStep(L, CycleNumber, dt, 0.0)) != ENZO_PPM_SUCCESS) break;
      if (hipStreamSynchronize(L->stream) != hipSuccess) { st = ENZO_PPM_ERR_HIP; break; }
      auto t0 = std::chrono::steady_clock::now();
      for (int r = 0; r < repeats && st == ENZO_PPM_SUCCESS; r++)
        st = hipPPMLevelStep(L, CycleNumber, dt, 0.0);
      if (hipStreamSynchronize(L->stream) != hipSuccess) st = ENZO_PPM_ERR_HIP;
      tp[pers] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (st == ENZO_PPM_SUCCESS) L->tune_times[4 * kShapes + pers] = tp[pers] / (double)repeats;
    }
    hipPPMLevelPersistent(L, tp[1] < tp[0] ? 1 : 0);
  }
  // restore the state and the counters
  L->parity = parity0;
  if (hipMemcpy(parity0 == 0 ? L->A : L->B, save, sizeof(double) * tot, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemset(L->counters, 0, 4 * sizeof(unsigned long long)) != hipSuccess)
    st = ENZO_PPM_ERR_HIP;
  (void)hipFree(save);
  if (shape_by_length)
    for (int c = 0; c < 4; c++) shape_by_length[c] = choice[c];
  return st;
}

int hipPPMLevelTuneTimes(const hipPPMLevel *L, enzo_float *times) {
  if (!L || !times) return ENZO_PPM_ERR_ARGUMENT;
  for (int k = 0; k < 4 * kShapes + 2; k++) times[k] = L->tune_times[k];
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelSetComoving(hipPPMLevel *L, enzo_int enable) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  if (enable) {
    if (!L->cws_dt) {
      if (dev_alloc(&L->cws_dt, L->hcw.size())) return ENZO_PPM_ERR_HIP;
      LCHK(hipMemcpy(L->cws_dt, L->hcw.data(), sizeof(double) * L->hcw.size(),
                     hipMemcpyHostToDevice));
    }
    if (!L->old) {
      long long tot = 0;
      L->ooff.assign(L->ngrids, 0);
      for (int g = 0; g < L->ngrids; g++) {
        L->ooff[g] = tot;
        tot += (long long)(S_GE + 1) * L->cells[g];
      }
      if (dev_alloc(&L->old, (size_t)tot)) return ENZO_PPM_ERR_HIP;
    }
  }
  L->comoving = enable != 0;
  L->old_valid = false;
  if (!L->comoving) L->a_sweep = L->aye = 1.0;
  int st = upload_geometry(L);
  if (st == ENZO_PPM_SUCCESS) st = upload_params(L);
  if (st == ENZO_PPM_SUCCESS) st = build_plan(L);  // graphs gain or lose the state copy
  return st;
}

int hipPPMLevelSetExpansionFactors(hipPPMLevel *L, enzo_float a_courant, enzo_float a_sweep) {
  if (!L || !L->comoving || !(a_courant > 0.0) || !(a_sweep > 0.0)) return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  L->aye = a_courant;
  int st = upload_params(L);
  if (st != ENZO_PPM_SUCCESS || a_sweep == L->a_sweep) return st;
  L->a_sweep = a_sweep;
  bool uni_changed = false;
  st = upload_geometry(L, &uni_changed);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (uni_changed) return build_plan(L);
  // same launches: only the uniform-axis constants in the device-resident
  // sweep parameters change, which captured graphs read at replay
  const int G = L->ngrids;
  for (int par = 0; par < 2; par++)
    for (int d = 0; d < L->rank; d++)
      for (int g = 0; g < G; g++) {
        SweepParams &S = L->hparams[((size_t)par * 3 + d) * G + g];
        const int rax = (d + 1) % 3, oax = (d + 2) % 3;
        for (int t = 0; t < NTAB; t++) {
          S.kt[t] = L->kt[((size_t)3 * g + d) * NTAB + t];
          S.kt_row[t] = L->kt[((size_t)3 * g + rax) * NTAB + t];
          S.kt_out[t] = L->kt[((size_t)3 * g + oax) * NTAB + t];
        }
        set_dx_inverse(S);
      }
  LCHK(hipMemcpy(L->dparams, L->hparams.data(), sizeof(SweepParams) * L->hparams.size(),
                 hipMemcpyHostToDevice));
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelTimeStep(hipPPMLevel *L, enzo_float cfl, enzo_float *dt_host) {
  if (!L || !dt_host) return ENZO_PPM_ERR_ARGUMENT;
  hipLaunchKernelGGL(level_dt_init_kernel, dim3(1), dim3(64), 0, L->stream, L->dtmin);
  hipLaunchKernelGGL(level_pressure_kernel, dim3((unsigned)L->npwork), dim3(256), 0, L->stream,
                     L->dpres + (size_t)L->parity * L->ngrids, L->dpwork);
  LCHK(hipGetLastError());
  hipLaunchKernelGGL(level_dt_final_kernel, dim3(1), dim3(64), 0, L->stream, L->dtmin, cfl, L->dt);
  LCHK(hipGetLastError());
  LCHK(hipMemcpyAsync(dt_host, L->dt, sizeof(double), hipMemcpyDeviceToHost, L->stream));
  LCHK(hipStreamSynchronize(L->stream));
  L->pres_current = true;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelComovingExpansionTerms(hipPPMLevel *L, enzo_float coef, enzo_float Coefficient) {
  if (!L || !L->comoving) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  LCHK(hipStreamSynchronize(L->stream));
  int st = upload_params(L, coef, Coefficient);
  if (st != ENZO_PPM_SUCCESS) return st;
  hipLaunchKernelGGL(level_expand_kernel, dim3((unsigned)L->npwork), dim3(256), 0, L->stream,
                     L->dexp + (size_t)L->parity * L->ngrids, L->dpwork);
  LCHK(hipGetLastError());
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelSetSiblings(hipPPMLevel *L, const enzo_long_int *GridGlobalStart,
                           const enzo_long_int *DomainActive, enzo_int periodic) {
  if (!L || !GridGlobalStart || !DomainActive) return ENZO_PPM_ERR_ARGUMENT;
  const int G = L->ngrids, rank = L->rank;
  std::vector<CopyBox> boxes;
  long long max_box = 0;
  const int sh_lo = periodic ? -1 : 0, sh_hi = periodic ? 1 : 0;
  for (int a = 0; a < G; a++)
    for (int b = 0; b < G; b++)
      for (int sz = (rank > 2 ? sh_lo : 0); sz <= (rank > 2 ? sh_hi : 0); sz++)
        for (int sy = (rank > 1 ? sh_lo : 0); sy <= (rank > 1 ? sh_hi : 0); sy++)
          for (int sx = sh_lo; sx <= sh_hi; sx++) {
            const int sft[3] = {sx, sy, sz};
            if (a == b && sx == 0 && sy == 0 && sz == 0) continue;
            CopyBox bx;
            bx.dst = a;
            bx.src = b;
            bool ok = true;
            for (int d = 0; d < 3 && ok; d++) {
              if (d >= rank) {
                bx.d0[d] = bx.s0[d] = 0;
                bx.n[d] = 1;
                continue;
              }
              const long long A0 = GridGlobalStart[3 * a + d] + sft[d] * DomainActive[d];
              const long long A1 = A0 + L->GD[3 * a + d];
              const long long B0 = GridGlobalStart[3 * b + d] + L->para[b].GridStartIndex[d];
              const long long B1 = GridGlobalStart[3 * b + d] + L->para[b].GridEndIndex[d] + 1;
              const long long lo = std::max(A0, B0), hi = std::min(A1, B1);
              if (hi <= lo) {
                ok = false;
                break;
              }
              bx.d0[d] = (int)(lo - A0);
              bx.s0[d] = (int)(lo - GridGlobalStart[3 * b + d]);
              bx.n[d] = (int)(hi - lo);
            }
            if (!ok) continue;
            boxes.push_back(bx);
            max_box = std::max(max_box, (long long)bx.n[0] * bx.n[1] * bx.n[2]);
          }
  if (boxes.size() > 65535) return ENZO_PPM_ERR_NOT_SUPPORTED;
  std::vector<GridGeom> geom(G);
  for (int g = 0; g < G; g++) {
    geom[g].off = L->off[g];
    geom[g].cells = L->cells[g];
    for (int d = 0; d < 3; d++) geom[g].GD[d] = L->GD[3 * g + d];
  }
  LCHK(hipStreamSynchronize(L->stream));
  if (L->dboxes) (void)hipFree(L->dboxes);
  L->dboxes = nullptr;
  if (!L->dgeom && dev_alloc(&L->dgeom, G)) return ENZO_PPM_ERR_HIP;
  if (dev_alloc(&L->dboxes, boxes.size())) return ENZO_PPM_ERR_HIP;
  if (!boxes.empty())
    LCHK(hipMemcpy(L->dboxes, boxes.data(), sizeof(CopyBox) * boxes.size(),
                   hipMemcpyHostToDevice));
  LCHK(hipMemcpy(L->dgeom, geom.data(), sizeof(GridGeom) * G, hipMemcpyHostToDevice));
  L->nboxes = (long long)boxes.size();
  L->max_box = max_box;
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelCopyZones(hipPPMLevel *L) {
  if (!L || !L->dgeom) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  if (L->nboxes == 0) return ENZO_PPM_SUCCESS;
  const unsigned bx =
      (unsigned)std::min<long long>((L->max_box * L->nslot + 255) / 256, 1024);
  hipLaunchKernelGGL(level_copy_kernel, dim3(bx, (unsigned)L->nboxes), dim3(256), 0, L->stream,
                     L->dboxes, L->dgeom, L->parity == 0 ? L->A : L->B, L->nslot);
  LCHK(hipGetLastError());
  return ENZO_PPM_SUCCESS;
}

int hipPPMLevelPeriodicBoundary(hipPPMLevel *L) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  const int G = L->ngrids;
  std::vector<BCParams> hb((size_t)3 * G);
  long long nmax = 1;
  for (int a = 0; a < L->rank; a++)
    for (int g = 0; g < G; g++) {
      const hipPPMParameter &P = L->para[g];
      BCParams &B = hb[(size_t)a * G + g];
      B.axis = a;
      B.ghost = (int)P.GridStartIndex[a];
      B.nslot = L->nslot;
      for (int d = 0; d < 3; d++) {
        B.GD[d] = L->GD[3 * g + d];
        B.act[d] = (int)(P.GridEndIndex[d] - P.GridStartIndex[d] + 1);
      }
      if (!periodic_fits(B, P.GridDimension[a] - P.GridEndIndex[a] - 1)) return ENZO_PPM_ERR_ARGUMENT;
      B.cells = L->cells[g];
      B.S = (L->parity == 0 ? L->A : L->B) + L->off[g];
      nmax = std::max(nmax, (long long)2 * B.ghost * B.GD[(a + 1) % 3] * B.GD[(a + 2) % 3] *
                                B.nslot);
    }
  BCParams *db = nullptr;
  if (dev_alloc(&db, hb.size())) return ENZO_PPM_ERR_HIP;
  int st = ENZO_PPM_SUCCESS;
  if (hipMemcpyAsync(db, hb.data(), sizeof(BCParams) * hb.size(), hipMemcpyHostToDevice,
                     L->stream) != hipSuccess)
    st = ENZO_PPM_ERR_HIP;
  const unsigned bx = (unsigned)std::min<long long>((nmax + 255) / 256, 65535);
  for (int a = 0; a < L->rank && st == ENZO_PPM_SUCCESS; a++) {
    hipLaunchKernelGGL(level_periodic_kernel, dim3(bx, (unsigned)G), dim3(256), 0, L->stream,
                       db + (size_t)a * G);
    if (hipGetLastError() != hipSuccess) st = ENZO_PPM_ERR_HIP;
  }
  if (hipStreamSynchronize(L->stream) != hipSuccess) st = ENZO_PPM_ERR_HIP;
  (void)hipFree(db);
  if (st != ENZO_PPM_SUCCESS) snprintf(g_level_error, sizeof(g_level_error), "periodic fill failed");
  return st;
}

int hipPPMLevelExternalBoundary(hipPPMLevel *L, const enzo_int *BoundaryType,
                                const enzo_int *GridOffset, const enzo_int *BoundaryDimension,
                                enzo_int NumberOfBaryonFields, const enzo_float *InflowValue) {
  if (!L || !BoundaryType || !GridOffset || !BoundaryDimension) return ENZO_PPM_ERR_ARGUMENT;
  L->pres_current = false;  // the state changes
  const int G = L->ngrids;
  // per face in the reference's order (x left, x right, y left, ...): the
  // grids that have it on the domain's edge; everything checked first
  std::vector<std::vector<ExtBCParams>> faces(6);
  for (int a = 0; a < L->rank; a++) {
    if (BoundaryDimension[a] <= 1) continue;
    for (int face = 0; face < 2; face++) {
      const int type = (int)BoundaryType[2 * a + face];
      if (type < BK_REFLECTING || type > BK_SHEARING) return ENZO_PPM_ERR_NOT_SUPPORTED;
      if (type == BK_PERIODIC) continue;
      for (int g = 0; g < G; g++) {
        const hipPPMParameter &P = L->para[g];
        const int *GD = &L->GD[3 * g];
        const bool touches = face == 0 ? GridOffset[3 * g + a] == 0
                                       : GridOffset[3 * g + a] + GD[a] == BoundaryDimension[a];
        const int S = (int)P.GridStartIndex[a], E = (int)P.GridEndIndex[a], n = E - S + 1;
        const int layers = face == 0 ? S : GD[a] - E - 1;
        if (!touches || layers <= 0) continue;
        if (type == BK_REFLECTING && (face == 0 ? 2 * S - 1 >= GD[a] : E - (layers - 1) < 0))
          return ENZO_PPM_ERR_ARGUMENT;
        if (type == BK_SHEARING && n < layers) return ENZO_PPM_ERR_NOT_SUPPORTED;
        ExtBCParams B;
        memset(&B, 0, sizeof(B));
        B.axis = a;
        B.face = face;
        B.type = type;
        B.layers = layers;
        B.S = S;
        B.E = E;
        B.n = n;
        for (int d = 0; d < 3; d++) B.GD[d] = GD[d];
        B.nslot = L->nslot;
        B.vslot = S_VX + a;
        B.cells = L->cells[g];
        B.A = (L->parity == 0 ? L->A : L->B) + L->off[g];
        if (type == BK_INFLOW) {
          if (!InflowValue) return ENZO_PPM_ERR_ARGUMENT;
          long long fnum[S_COL0 + ENZO_PPM_MAX_COLOR];
          for (int s = 0; s < L->nslot; s++) fnum[s] = -1;
          fnum[S_D] = P.DensNum;
          fnum[S_E] = P.TENum;
          fnum[S_VX] = P.Vel1Num;
          if (P.GridRank > 1) fnum[S_VY] = P.Vel2Num;
          if (P.GridRank > 2) fnum[S_VZ] = P.Vel3Num;
          if (P.DualEnergyFormalism) fnum[S_GE] = P.GENum;
          for (int c = 0; c < L->ncol; c++) fnum[S_COL0 + c] = P.colnum[c];
          for (int s = 0; s < L->nslot; s++) {
            if (fnum[s] >= NumberOfBaryonFields) return ENZO_PPM_ERR_ARGUMENT;
            B.inflow[s] = fnum[s] >= 0 ? InflowValue[fnum[s] * 6 + 2 * a + face] : 0.0;
          }
        }
        faces[2 * a + face].push_back(B);
      }
    }
  }
  size_t tot = 0;
  for (auto &v : faces) tot += v.size();
  if (tot == 0) return ENZO_PPM_SUCCESS;
  std::vector<ExtBCParams> hb;
  hb.reserve(tot);
  for (auto &v : faces) hb.insert(hb.end(), v.begin(), v.end());
  ExtBCParams *db = nullptr;
  if (dev_alloc(&db, hb.size())) return ENZO_PPM_ERR_HIP;
  int st = ENZO_PPM_SUCCESS;
  if (hipMemcpyAsync(db, hb.data(), sizeof(ExtBCParams) * hb.size(), hipMemcpyHostToDevice,
                     L->stream) != hipSuccess)
    st = ENZO_PPM_ERR_HIP;
  size_t first = 0;
  for (int f = 0; f < 6 && st == ENZO_PPM_SUCCESS; f++) {
    const size_t cnt = faces[f].size();
    if (cnt == 0) continue;
    long long nmax = 1;
    for (const auto &B : faces[f])
      nmax = std::max(nmax, (long long)B.layers * B.GD[(B.axis + 1) % 3] * B.GD[(B.axis + 2) % 3] *
                                B.nslot);
    const unsigned bx = (unsigned)std::min<long long>((nmax + 255) / 256, 65535);
    hipLaunchKernelGGL(level_ext_bc_kernel, dim3(bx, (unsigned)cnt), dim3(256), 0, L->stream,
                       db + first);
    if (hipGetLastError() != hipSuccess) st = ENZO_PPM_ERR_HIP;
    first += cnt;
  }
  if (hipStreamSynchronize(L->stream) != hipSuccess) st = ENZO_PPM_ERR_HIP;
  (void)hipFree(db);
  if (st != ENZO_PPM_SUCCESS) snprintf(g_level_error, sizeof(g_level_error), "external boundary failed");
  return st;
}

int hipPPMLevelSetLearnedGuess(hipPPMLevel *L, const float *weights) {
  if (!L) return ENZO_PPM_ERR_ARGUMENT;
  LCHK(hipStreamSynchronize(L->stream));
  if (weights && !learn_weights_valid(weights)) return ENZO_PPM_ERR_ARGUMENT;
  // a closed gate (no interface can pass) runs the plain kernels
  const bool on = weights && !learn_gate_closed(weights);
  if (on) {
    if (!L->learn && dev_alloc(&L->learn, kMlpDevFloats)) return ENZO_PPM_ERR_HIP;
    float dw[kMlpDevFloats];
    learn_device_weights(weights, dw);
    LCHK(hipMemcpy(L->learn, dw, sizeof(float) * kMlpDevFloats, hipMemcpyHostToDevice));
  }
  L->learned = on;
  return build_plan(L);  // parameters carry the network pointer; graphs are recaptured
}

int hipPPMLevelDestroy(hipPPMLevel *L) {
  if (!L) return ENZO_PPM_SUCCESS;
  free_level(L);
  delete L;
  return ENZO_PPM_SUCCESS;
}

const char *hipPPMLevelLastError(void) { return g_level_error; }

}  // extern "C"
