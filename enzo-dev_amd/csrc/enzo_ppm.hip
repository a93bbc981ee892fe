// enzo_ppm.hip -- MI355X (gfx950) implementation of the C ABI in
// include/enzo_ppm.h: device state, kernels and the cuPPM-style lifecycle.
//
// Device state: the baryon fields live in two buffers of NSLOT arrays each
// (slots: d, e, vx, vy, vz, ge, colours), x fastest, ghost zones included.
// A sweep reads buffer A and writes buffer B (tiles read halo cells owned by
// neighbouring tiles, so the update cannot be in place); the buffers swap
// after each sweep.  Pressure, acceleration and per-axis coefficient tables
// are separate arrays.  Nothing is re-allocated while the grid shape fits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/enzo_ppm.h"
#include "ppm_device.hpp"
#include "ppm_geometry.hpp"
#include "ppm_stage.hpp"
#include "ppm_sweep.hpp"

using namespace ppm;

namespace {

constexpr int kTP = kGridTP;  // pencils per workgroup
constexpr int kTS = kGridTS;  // interfaces per pencil segment
using SweepCfg = Cfg<kTP, kTS>;

// ------------------------------------------------------------------------
// Fused sweep kernel: see ppm_sweep.hpp for the stages.
template <int UNI, int LEARN>
// 5 wavefronts per SIMD: 96 VGPRs (a few spilled) and 31 KB of LDS per
// workgroup, 5 workgroups per CU
__global__ __launch_bounds__(kTP *kTS) __attribute__((amdgpu_waves_per_eu(5)))
void ppm_sweep_kernel(SweepParams P) {
  __shared__ double sm[SweepCfg::SMEM];
  PPM_LDS_ENTER(sm);
  DevExec ex;
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  TileCounts tc = sweep_tile<kTP, kTS, UNI, DevExec, LEARN>(P, tile, sm, ex);
  if (threadIdx.x == 0) {
    // a failing cell or an error: the fallback kernel redoes the group's
    // pencils whole (the fallback chain can cross segments) and settles them
    if (tc.bad || tc.err) flag_group(P, tile / P.nseg);
    if (tc.chains) atomicAdd(&P.counters[1], (unsigned long long)tc.chains);
  }
}

// ------------------------------------------------------------------------
// Per-cell HLL fallback (flux_twoshock.F:401-456): every pencil of each
// group the sweep kernel flagged is recomputed whole by one workgroup of
// kFbTS threads (one thread per interface, the pencil in LDS), with the
// fallback applied in the reference's cell order; its results replace the
// sweep kernel's for those pencils.  Launched after every sweep launch; with
// no flagged group each workgroup only reads the count.
template <int UNI, int LEARN, int TS = kFbTS>
__global__ __launch_bounds__(TS) void ppm_fallback_kernel(SweepParams P) {
  __shared__ double sm[NARR_FB * Cfg<1, TS>::LW];
  PPM_LDS_ENTER(sm);
  DevExec ex;
  const unsigned n = *P.fb_count;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x)
    fallback_group<UNI, LEARN, TS>(P, (int)P.fb_list[k], sm, ex);
  fallback_launch_done(P.fb_count, P.fb_done);
}

// ------------------------------------------------------------------------
// RiemannSolver HLL: every pencil group of the launch on the fallback list
// (the one-pencil workgroups then take flux_hll's fluxes everywhere).
__global__ void fb_all_kernel(SweepParams P, unsigned ngroups) {
  for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < ngroups; k += gridDim.x * blockDim.x)
    P.fb_list[k] = k;
  if (blockIdx.x == 0 && threadIdx.x == 0) *P.fb_count = ngroups;
}

// ------------------------------------------------------------------------
// Per-sweep preparation (prep_item in ppm_device.hpp), one item per thread.
__global__ void sweep_prep_kernel(FlatParams F) {
  prep_item(F, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------------
// Fused pressure + Courant condition, grid-stride over the cells.
__global__ __launch_bounds__(256) void pressure_dt_kernel(PresParams Q) {
  double local = 1.0e20;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < Q.cells;
       g += (long long)gridDim.x * blockDim.x)
    local = fmn(local, pressure_dt_cell(Q, g));
  if (Q.want_dt) block_min_dt(local, Q.dtmin);
}

// projection of one face's fine fluxes: item = (array, coarse position);
// each sums its fine cells in the reference's loop order (k, j, i)
struct ProjParams {
  const double *fine;  // [array][fine position]
  double *coarse;      // [array][coarse position]
  long long nfine, ncoarse;
  int narrays;
  int Dims[3], PD[3], r[3];
  double dArea;
};

__global__ void project_kernel(ProjParams Q) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)Q.narrays * Q.ncoarse) return;
  const int a = (int)(q / Q.ncoarse);
  const long long p = q % Q.ncoarse;
  const int i1 = (int)(p % Q.PD[0]);
  const int j1 = (int)((p / Q.PD[0]) % Q.PD[1]);
  const int k1 = (int)(p / ((long long)Q.PD[0] * Q.PD[1]));
  const double *f = Q.fine + (long long)a * Q.nfine;
  double acc = 0.0;
  for (int k = k1 * Q.r[2]; k < (k1 + 1) * Q.r[2] && k < Q.Dims[2]; k++)
    for (int j = j1 * Q.r[1]; j < (j1 + 1) * Q.r[1] && j < Q.Dims[1]; j++)
      for (int i = i1 * Q.r[0]; i < (i1 + 1) * Q.r[0] && i < Q.Dims[0]; i++)
        acc += f[i + (long long)j * Q.Dims[0] + (long long)k * Q.Dims[0] * Q.Dims[1]] * Q.dArea;
  Q.coarse[(long long)a * Q.ncoarse + p] = acc;
}

__global__ void fluxcorr_kernel(FluxCorrParams F) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= F.nface) return;
  const int i = (int)(p % F.Dim[0]);
  const int j = (int)((p / F.Dim[0]) % F.Dim[1]);
  const int k = (int)(p / ((long long)F.Dim[0] * F.Dim[1]));
  const long long x =
      ((long long)(F.Start[2] + k) * F.GD[1] + (F.Start[1] + j)) * F.GD[0] + (F.Start[0] + i);
  fluxcorr_side(F, x, 0, p);
  fluxcorr_side(F, x + F.Offset, 1, p);
}

__global__ void expand_kernel(ExpandParams X) {
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < X.cells;
       g += (long long)gridDim.x * blockDim.x)
    expand_cell(X, g);
}

__global__ void dt_init_kernel(unsigned long long *dtmin) {
  if (threadIdx.x == 0) *dtmin = (unsigned long long)__double_as_longlong(1.0e20);
}

__global__ void dt_final_kernel(const unsigned long long *dtmin, double cfl, double *dt) {
  if (threadIdx.x == 0) *dt = __longlong_as_double((long long)*dtmin) * cfl;
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// zero fill of n 32-bit words (used where a caller may capture the stream,
// instead of hipMemsetAsync)
__global__ void zero_u32_kernel(unsigned *p, long long n) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = 0u;
}

// the caller's separate field arrays <-> the slot-contiguous state
struct SlotPtrs {
  double *f[S_COL0 + ENZO_PPM_MAX_COLOR];  // per slot, or null (gather: zeros)
};
__global__ void gather_slots_kernel(double *A, long long cells, int nslot, SlotPtrs F) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nslot * cells;
       k += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(k / cells);
    const double *src = F.f[s];
    A[k] = src ? src[k - s * cells] : 0.0;
  }
}
__global__ void scatter_slots_kernel(const double *A, long long cells, int nslot, SlotPtrs F) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nslot * cells;
       k += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(k / cells);
    if (F.f[s]) F.f[s][k - s * cells] = A[k];
  }
}

// ------------------------------------------------------------------------
// Periodic ghost fill along one axis for every slot (the periodic case of
// the external boundary).  Thread per (slot, ghost cell).
__global__ void periodic_axis_kernel(BCParams B) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < periodic_items(B)) periodic_item(B, (int)gid);
}

__global__ void ext_bc_kernel(ExtBCParams B) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < ext_bc_items(B)) ext_bc_item(B, gid);
}

}  // namespace

// ==========================================================================
// host side
struct hipPPMData {
  int device = 0;
  hipStream_t stream = nullptr;
  int GD[3] = {0, 0, 0};
  long long cells = 0;
  int nslot = 0, ncol = 0;
  bool gravity = false;
  double *A = nullptr, *B = nullptr;
  double *pres = nullptr;
  double *accel = nullptr;  // 3 * cells when gravity
  const double *accel_ext[3] = {nullptr, nullptr, nullptr};  // the caller's (no-alloc entry)
  double *tab[3] = {nullptr, nullptr, nullptr};
  double *cwdev[3] = {nullptr, nullptr, nullptr};
  double *flat = nullptr;
  long long flat_cap = 0;
  double *qc = nullptr;  // dt/dx of the current sweep axis
  double *dt = nullptr;
  unsigned long long *dtmin = nullptr;
  unsigned long long *counters = nullptr;  // [0] fallback [1] DE chain [2] errors [3] unsupported
  // per-cell fallback: group flags, flagged-group list, {count, done}
  int *fb_flag = nullptr;
  unsigned *fb_list = nullptr, *fb_state = nullptr;
  long long fb_groups = 0;
  int *segbuf = nullptr;  // per axis: o[nseg], w[nseg]
  int uni[3] = {0, 0, 0};  // axis has one cell width
  double kt[3][NTAB];      // its coefficient entries
  int segoff[3] = {0, 0, 0}, nseg[3] = {0, 0, 0};
  // face capture
  int *face_head[3] = {nullptr, nullptr, nullptr};
  FaceRec *faces[3] = {nullptr, nullptr, nullptr};
  double *face_out[3] = {nullptr, nullptr, nullptr};
  std::vector<FaceRec> hfaces[3];
  std::vector<int> hface_sub[3], hface_side[3];
  long long face_total[3] = {0, 0, 0};
  int nsub = 0;
  long long fallback_total = 0;
  long long chain_total = 0;
  // learned two-shock start (network parameters in the device layout,
  // kMlpDevFloats floats) when set
  float *learn = nullptr;
  bool learned = false;
  // comoving coordinates: sweeps use widths a_sweep * CellWidth, the Courant
  // pass CellWidth (cwdt) with aye; the state at each step's start is kept
  // in `old` (slots S_D..S_GE) for the time-centred expansion pressure
  bool comoving = false;
  double a_sweep = 1.0, aye = 1.0;
  double *cwdt[3] = {nullptr, nullptr, nullptr};
  double *old = nullptr;
  bool old_valid = false;
  // harvest of interface states (SweepParams::harvest) when harvest_cap > 0
  double *harvest = nullptr;
  unsigned long long *harvest_n = nullptr;
  long long harvest_cap = 0;
  unsigned harvest_keep = 1;
  // optional per-launch timing of the sweep kernel (HIP events on D->stream)
  int profile = 0;
  std::vector<hipEvent_t> ev;  // pairs
  size_t ev_used = 0;
  // pinned staging ring of the host-array transfers (stage_upload /
  // stage_download), allocated on first use; stage_busy[b]: a copy into or
  // out of buffer b was queued and stage_ev[b] marks its end
  void *stage[3] = {nullptr, nullptr, nullptr};
  hipEvent_t stage_ev[3] = {nullptr, nullptr, nullptr};
  bool stage_busy[3] = {false, false, false};
  int stage_next = 0;
};

namespace {

thread_local char g_last_error[512] = "";

void set_error(const char *what, hipError_t e, int line) {
  snprintf(g_last_error, sizeof(g_last_error), "%s failed: %s (enzo_ppm.hip:%d)", what,
           hipGetErrorString(e), line);
  if (getenv("ENZO_PPM_VERBOSE")) fprintf(stderr, "enzo_ppm: %s\n", g_last_error);
}

#define HIPCHK(x)                            \
  do {                                       \
    hipError_t e_ = (x);                     \
    if (e_ != hipSuccess) {                  \
      set_error(#x, e_, __LINE__);           \
      return ENZO_PPM_ERR_HIP;               \
    }                                        \
  } while (0)

inline int nslot_of(const hipPPMParameter *P) { return S_COL0 + (int)P->NumberOfColours; }

int check_para(const hipPPMParameter *P) {
  if (!P) return ENZO_PPM_ERR_ARGUMENT;
  if (P->GridRank < 1 || P->GridRank > 3) return ENZO_PPM_ERR_ARGUMENT;
  if (P->NumberOfColours < 0 || P->NumberOfColours > ENZO_PPM_MAX_COLOR) return ENZO_PPM_ERR_ARGUMENT;
  for (int d = 0; d < 3; d++) {
    if (P->GridDimension[d] < 1 || P->GridDimension[d] > ENZO_PPM_MAX_ANY_SINGLE_DIRECTION)
      return ENZO_PPM_ERR_ARGUMENT;
    if (d < P->GridRank) {
      if (P->GridStartIndex[d] < 3 || P->GridEndIndex[d] + 3 >= P->GridDimension[d] ||
          P->GridEndIndex[d] < P->GridStartIndex[d])
        return ENZO_PPM_ERR_ARGUMENT;
    }
  }
  if (P->ConservativeReconstruction != 0 || P->PositiveReconstruction != 0)
    return ENZO_PPM_ERR_NOT_SUPPORTED;  // off on this path (inteuler.F:232-239)
  if (P->RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK && P->RiemannSolver != ENZO_PPM_RIEMANN_HLL &&
      P->RiemannSolver != ENZO_PPM_RIEMANN_HLLC)
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  // the HLL and HLLC solvers run in one-pencil workgroups: every swept
  // pencil must fit
  if (P->RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK)
    for (int d = 0; d < P->GridRank; d++)
      if (P->GridDimension[d] > kFbTS + 1) {
        snprintf(g_last_error, sizeof(g_last_error),
                 "RiemannSolver HLL/HLLC: pencils longer than %d cells are not supported", kFbTS + 1);
        return ENZO_PPM_ERR_NOT_SUPPORTED;
      }
  if (P->PPMFlatteningParameter < 0 || P->PPMFlatteningParameter > 3)
    return ENZO_PPM_ERR_NOT_SUPPORTED;  // calcdiss.F has types 1-3
  if (P->PPMDiffusionParameter != 0 && P->PPMDiffusionParameter != 1)
    return ENZO_PPM_ERR_NOT_SUPPORTED;  // idiff = 2 reads undefined values in calcdiss.F
  return ENZO_PPM_SUCCESS;
}

// true when the allocation failed (so that `if (alloc_dev(..) || ..)` reads
// as "if any allocation failed")
bool alloc_dev(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc", e, __LINE__);
    *p = nullptr;
    return true;
  }
  return false;
}

void free_data(hipPPMData *D) {
  if (!D) return;
  for (hipEvent_t e : D->ev) (void)hipEventDestroy(e);
  D->ev.clear();
  void *ptrs[] = {D->A, D->B, D->pres, D->accel, D->tab[0], D->tab[1], D->tab[2], D->cwdev[0],
                  D->cwdev[1], D->cwdev[2], D->flat, D->qc, D->dt, D->dtmin, D->counters, D->segbuf,
                  D->fb_flag, D->fb_list, D->fb_state, D->learn, D->harvest, D->harvest_n,
                  D->cwdt[0], D->cwdt[1], D->cwdt[2], D->old,
                  D->face_head[0], D->face_head[1], D->face_head[2], D->faces[0], D->faces[1],
                  D->faces[2], D->face_out[0], D->face_out[1], D->face_out[2]};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int b = 0; b < 3; b++) {
    if (D->stage_ev[b]) (void)hipEventDestroy(D->stage_ev[b]);
    if (D->stage[b]) (void)hipHostFree(D->stage[b]);
    D->stage_ev[b] = nullptr;
    D->stage[b] = nullptr;
  }
}

// ------------------------------------------------------------------------
// Host-array transfers.  hipMemcpy from the caller's pageable arrays moves
// ~28 GB/s to the device and ~18 GB/s back (the runtime stages them
// itself); here a field goes in chunks through a ring of pinned buffers,
// the host copy of each chunk (CopyPool, several threads) overlapping the
// DMA of the previous one, which runs at ~57 GB/s.  Smaller transfers, or a
// host that cannot pin the ring, take the direct path.
constexpr int kStageBufs = 3;
constexpr size_t kStageBytes = size_t(32) << 20;  // per pinned buffer
constexpr size_t kStageMin = size_t(8) << 20;     // direct path below this total

struct HostXfer {
  char *dst;
  const char *src;
  size_t bytes;
};

bool stage_ready(hipPPMData *D) {
  if (D->stage[kStageBufs - 1]) return true;
  for (int b = 0; b < kStageBufs; b++) {
    if (!D->stage[b] && hipHostMalloc(&D->stage[b], kStageBytes, hipHostMallocDefault) != hipSuccess) {
      D->stage[b] = nullptr;
      return false;
    }
    if (!D->stage_ev[b] && hipEventCreateWithFlags(&D->stage_ev[b], hipEventDisableTiming) != hipSuccess) {
      D->stage_ev[b] = nullptr;
      (void)hipHostFree(D->stage[b]);
      D->stage[b] = nullptr;
      return false;
    }
  }
  return true;
}

size_t xfer_total(const std::vector<HostXfer> &xs) {
  size_t t = 0;
  for (const HostXfer &x : xs) t += x.bytes;
  return t;
}

// host -> device on D->stream; the caller's arrays may be reused on return
// (the DMA reads only the pinned ring)
int stage_upload(hipPPMData *D, const std::vector<HostXfer> &xs) {
  if (xfer_total(xs) < kStageMin || !stage_ready(D)) {
    for (const HostXfer &x : xs)
      HIPCHK(hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyHostToDevice, D->stream));
    return ENZO_PPM_SUCCESS;
  }
  int b = D->stage_next;
  for (const HostXfer &x : xs)
    for (size_t off = 0; off < x.bytes; off += kStageBytes) {
      const size_t len = std::min(kStageBytes, x.bytes - off);
      if (D->stage_busy[b]) HIPCHK(hipEventSynchronize(D->stage_ev[b]));
      CopyPool::get().copy(D->stage[b], x.src + off, len);
      HIPCHK(hipMemcpyAsync(x.dst + off, D->stage[b], len, hipMemcpyHostToDevice, D->stream));
      HIPCHK(hipEventRecord(D->stage_ev[b], D->stream));
      D->stage_busy[b] = true;
      b = (b + 1) % kStageBufs;
    }
  D->stage_next = b;
  return ENZO_PPM_SUCCESS;
}

// device -> host after the work queued on D->stream; complete on return
int stage_download(hipPPMData *D, const std::vector<HostXfer> &xs) {
  if (xfer_total(xs) < kStageMin || !stage_ready(D)) {
    for (const HostXfer &x : xs)
      HIPCHK(hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyDeviceToHost, D->stream));
    HIPCHK(hipStreamSynchronize(D->stream));
    return ENZO_PPM_SUCCESS;
  }
  char *pend_dst[kStageBufs] = {};
  size_t pend_len[kStageBufs] = {};
  auto drain = [&](int c) -> int {  // buffer c's chunk to the caller's array
    if (!pend_dst[c]) return ENZO_PPM_SUCCESS;
    HIPCHK(hipEventSynchronize(D->stage_ev[c]));
    CopyPool::get().copy(pend_dst[c], D->stage[c], pend_len[c]);
    pend_dst[c] = nullptr;
    D->stage_busy[c] = false;
    return ENZO_PPM_SUCCESS;
  };
  int b = D->stage_next;
  for (const HostXfer &x : xs)
    for (size_t off = 0; off < x.bytes; off += kStageBytes) {
      const size_t len = std::min(kStageBytes, x.bytes - off);
      if (pend_dst[b]) {
        const int st = drain(b);
        if (st != ENZO_PPM_SUCCESS) return st;
      } else if (D->stage_busy[b]) {
        HIPCHK(hipEventSynchronize(D->stage_ev[b]));
      }
      HIPCHK(hipMemcpyAsync(D->stage[b], x.src + off, len, hipMemcpyDeviceToHost, D->stream));
      HIPCHK(hipEventRecord(D->stage_ev[b], D->stream));
      D->stage_busy[b] = true;
      pend_dst[b] = x.dst + off;
      pend_len[b] = len;
      b = (b + 1) % kStageBufs;
    }
  for (int i = 0; i < kStageBufs; i++) {
    const int st = drain((b + i) % kStageBufs);
    if (st != ENZO_PPM_SUCCESS) return st;
  }
  D->stage_next = b;
  return ENZO_PPM_SUCCESS;
}

int launch_sweep(hipPPMData *D, const hipPPMParameter *P, double dtFixed, const double *dtp,
                 int dim) {
  const int N = D->GD[dim];
  const int i1 = (int)P->GridStartIndex[dim] + 1, i2 = (int)P->GridEndIndex[dim] + 1;
  const int rax = (dim + 1) % 3, oax = (dim + 2) % 3;
  const int rank = (int)P->GridRank;
  const int iflat = (int)P->PPMFlatteningParameter;
  // 1) dt/dx table and flattening row of line 1 of every slice
  {
    FlatParams F;
    F.dim = dim; F.N = N; F.i1 = i1; F.i2 = i2;
    for (int d = 0; d < 3; d++) F.GD[d] = D->GD[d];
    F.cells = D->cells;
    F.n_outer = D->GD[oax];
    F.iflatten = iflat;
    F.gamma = P->Gamma;
    F.A = D->A;
    F.pres = D->pres;
    F.flat = D->flat;
    F.dx = D->cwdev[dim];
    F.dt = dtFixed;
    F.dtp = dtp;
    F.qc = D->qc;
    long long n = N + (iflat ? (long long)(i2 - i1 + 3) * F.n_outer : 0);
    hipLaunchKernelGGL(sweep_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       D->stream, F);
    HIPCHK(hipGetLastError());
  }
  SweepParams S;
  memset(&S, 0, sizeof(S));
  S.dim = dim; S.N = N; S.i1 = i1; S.i2 = i2; S.rank = rank;
  for (int d = 0; d < 3; d++) S.GD[d] = D->GD[d];
  S.cells = D->cells;
  S.nseg = D->nseg[dim];
  S.seg_o = D->segbuf + D->segoff[dim];
  S.seg_w = D->segbuf + D->segoff[dim] + D->nseg[dim];
  S.A = D->A;
  S.B = D->B;
  S.ncol = (int)P->NumberOfColours;
  S.pres = D->pres;
  S.gr = !P->GravityOn ? nullptr
         : D->accel_ext[dim] ? D->accel_ext[dim] : D->accel + (long long)dim * D->cells;
  S.flat = D->flat;
  S.tab = D->tab[dim];
  S.tab_row = D->tab[rax];
  S.tab_out = D->tab[oax];
  S.qc = D->qc;
  S.uni[0] = D->uni[dim];
  S.uni[1] = D->uni[rax];
  S.uni[2] = D->uni[oax];
  for (int t = 0; t < NTAB; t++) {
    S.kt[t] = D->kt[dim][t];
    S.kt_row[t] = D->kt[rax][t];
    S.kt_out[t] = D->kt[oax][t];
  }
  set_dx_inverse(S);
  S.gamma = P->Gamma;
  S.eta1 = P->DualEnergyFormalismEta1;
  S.eta2 = P->DualEnergyFormalismEta2;
  S.pmin = P->MinimumPressure;
  S.dfloor = P->SmallRho;
  S.dt = dtFixed;
  S.dtp = dtp;
  S.isteep = (int)P->PPMSteepeningParameter;
  S.iflatten = iflat;
  S.idiff = (int)P->PPMDiffusionParameter;
  S.idual = (int)P->DualEnergyFormalism;
  S.ipresfree = (int)P->PressureFree;
  S.gravity = (int)P->GravityOn;
  S.ifallback = (int)P->RiemannSolverFallback;
  S.nact_out = (int)(P->GridEndIndex[oax] - P->GridStartIndex[oax] + 1);
  S.face_head = D->face_total[dim] > 0 ? D->face_head[dim] : nullptr;
  S.faces = D->faces[dim];
  S.face_out = D->face_out[dim];
  S.counters = D->counters;
  S.fb_flag = D->fb_flag;
  S.fb_list = D->fb_list;
  S.fb_count = D->fb_state;
  S.fb_done = D->fb_state + 1;
  S.fb_tag = 0;
  S.tp = kTP;
  S.fb_ok = pencil_fits<kFbTS>(S) ? 1 : 0;
  S.hll_all = P->RiemannSolver == ENZO_PPM_RIEMANN_HLL ? 1 : 0;
  S.hllc = P->RiemannSolver == ENZO_PPM_RIEMANN_HLLC ? 1 : 0;
  S.learn = D->learned ? D->learn : nullptr;
  if (D->harvest_cap > 0) {
    S.harvest = D->harvest;
    S.harvest_n = D->harvest_n;
    S.harvest_cap = D->harvest_cap;
    S.harvest_keep = D->harvest_keep;
  }
  // instantiation by (one cell width, learned start)
  const int variant = (S.uni[0] ? 1 : 0) | (D->learned ? 2 : 0);
  auto launch_main = [&](unsigned nblk) {
    const dim3 g(nblk), b(kTP * kTS);
    switch (variant) {
      case 0: hipLaunchKernelGGL((ppm_sweep_kernel<0, 0>), g, b, 0, D->stream, S); break;
      case 1: hipLaunchKernelGGL((ppm_sweep_kernel<1, 0>), g, b, 0, D->stream, S); break;
      case 2: hipLaunchKernelGGL((ppm_sweep_kernel<0, 1>), g, b, 0, D->stream, S); break;
      default: hipLaunchKernelGGL((ppm_sweep_kernel<1, 1>), g, b, 0, D->stream, S); break;
    }
    return hipGetLastError();
  };
  // the two-shock tiles, or with the HLL solver every group to the
  // one-pencil workgroups (which then do the whole sweep)
  const bool whole_pencils = S.hll_all || S.hllc;
  auto launch_main_or_all = [&](unsigned nblk) {
    if (!whole_pencils) return launch_main(nblk);
    const unsigned ngroups = nblk / (unsigned)S.nseg;
    hipLaunchKernelGGL(fb_all_kernel, dim3((ngroups + 255) / 256), dim3(256), 0, D->stream, S,
                       ngroups);
    return hipGetLastError();
  };
  // one-pencil workgroups: the narrowest width that holds the pencils
  // (fb_ok above: the widest does)
  const int fbts = S.fb_ok ? fb_width_of(S) : kFbTS;
  // ngroups: pencil groups of the launch, the most that can be on the list
  // (any number of workgroups is correct: each takes every gridDim-th entry)
  auto launch_fb = [&](unsigned ngroups) {
    const unsigned nblk = std::max(1u, std::min(ngroups, fb_blocks(fbts, (whole_pencils ? kFbBlocksWhole : kFbBlocks) * kFbTS)));
    return with_fb_width(fbts, [&](auto W) {
      constexpr int TS = decltype(W)::value;
      const dim3 g(nblk), b(TS);
      switch (variant) {
        case 0: hipLaunchKernelGGL((ppm_fallback_kernel<0, 0, TS>), g, b, 0, D->stream, S); break;
        case 1: hipLaunchKernelGGL((ppm_fallback_kernel<1, 0, TS>), g, b, 0, D->stream, S); break;
        case 2: hipLaunchKernelGGL((ppm_fallback_kernel<0, 1, TS>), g, b, 0, D->stream, S); break;
        default: hipLaunchKernelGGL((ppm_fallback_kernel<1, 1, TS>), g, b, 0, D->stream, S); break;
      }
    });
  };
  if (S.idiff == 0) {
    // all pencils at once: t1 is the contiguous transverse axis for y/z
    // sweeps (coalesced loads), the row axis for x sweeps
    if (dim == 0) { S.ax_t1 = 1; S.ax_t2 = 2; }
    else if (dim == 1) { S.ax_t1 = 0; S.ax_t2 = 2; }
    else { S.ax_t1 = 0; S.ax_t2 = 1; }
    S.n_t1 = D->GD[S.ax_t1];
    S.n_t2 = D->GD[S.ax_t2];
    S.t2_0 = 0;
    set_strides(S);
    S.nblk_t1 = (S.n_t1 + kTP - 1) / kTP;
    long long nblk = (long long)S.nseg * S.nblk_t1 * S.n_t2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (D->profile) {
      if (D->ev_used + 2 > D->ev.size()) {
        for (int k = 0; k < 2; k++) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          D->ev.push_back(e);
        }
      }
      e0 = D->ev[D->ev_used];
      e1 = D->ev[D->ev_used + 1];
      D->ev_used += 2;
      HIPCHK(hipEventRecord(e0, D->stream));
    }
    HIPCHK(launch_main_or_all((unsigned)nblk));
    if (whole_pencils) HIPCHK(launch_fb((unsigned)(nblk / S.nseg)));
    if (D->profile) HIPCHK(hipEventRecord(e1, D->stream));
    if (!whole_pencils) HIPCHK(launch_fb((unsigned)(nblk / S.nseg)));
  } else {
    // diffusion reads the live transverse velocities of the previous slice
    // (calcdiss.F:207-288 on BaryonField): slices in order, one launch each
    S.ax_t1 = rax;
    S.ax_t2 = oax;
    S.n_t1 = D->GD[rax];
    S.n_t2 = 1;
    set_strides(S);
    S.nblk_t1 = (S.n_t1 + kTP - 1) / kTP;
    for (int k = 0; k < D->GD[oax]; k++) {
      S.t2_0 = k;
      long long nblk = (long long)S.nseg * S.nblk_t1;
      HIPCHK(launch_main_or_all((unsigned)nblk));
      HIPCHK(launch_fb((unsigned)S.nblk_t1));  // the next slice reads this one's results
    }
  }
  std::swap(D->A, D->B);
  return ENZO_PPM_SUCCESS;
}

// cell widths of the sweeps (a_sweep * CellWidth in comoving coordinates,
// Grid_SolveHydroEquations.C:386-408, the product taken as the reference
// takes it) and their coefficient tables; the Courant pass's widths
// (CellWidth itself) when comoving
int upload_geometry(hipPPMData *D, const hipPPMParameter *Para) {
  // host staging of one axis (no heap allocation: the entry point that
  // allocates nothing uses this too)
  constexpr int kN = ENZO_PPM_MAX_ANY_SINGLE_DIRECTION;
  static thread_local double cw[kN], dx[kN], tab[(size_t)NTAB * kN];
  for (int d = 0; d < 3; d++) {
    const int N = D->GD[d];
    for (int i = 0; i < N; i++)
      cw[i] = Para->CellWidth[d] ? Para->CellWidth[d][i] : (d < Para->GridRank ? Para->dx[d] : 1.0);
    for (int i = 0; i < N; i++)
      dx[i] = D->comoving && d < Para->GridRank ? D->a_sweep * cw[i] : cw[i];
    build_tables(dx, N, tab);
    D->uni[d] = uniform_axis(dx, N, tab, D->kt[d]) ? 1 : 0;
    HIPCHK(hipMemcpyAsync(D->tab[d], tab, sizeof(double) * NTAB * (size_t)N,
                          hipMemcpyHostToDevice, D->stream));
    HIPCHK(hipMemcpyAsync(D->cwdev[d], dx, sizeof(double) * N, hipMemcpyHostToDevice,
                          D->stream));
    if (D->comoving)
      HIPCHK(hipMemcpyAsync(D->cwdt[d], cw, sizeof(double) * N, hipMemcpyHostToDevice,
                            D->stream));
    HIPCHK(hipStreamSynchronize(D->stream));  // the staging is reused for the next axis
  }
  return ENZO_PPM_SUCCESS;
}

// Sweep segments of every axis into segs[] (o[nseg] then w[nseg] per axis,
// D->segoff / D->nseg set); returns the number of ints or -1 when an axis
// cannot be split.  segs holds 2 * 3 * ENZO_PPM_MAX_ANY_SINGLE_DIRECTION.
int plan_segments(hipPPMData *D, const hipPPMParameter *Para, int *segs) {
  int n = 0;
  for (int d = 0; d < 3; d++) {
    const int N = D->GD[d];
    const int i1 = (int)Para->GridStartIndex[d] + 1, i2 = (int)Para->GridEndIndex[d] + 1;
    int ns = 0;
    if (d < Para->GridRank && N > 1) {
      ns = build_segments<kTP, kTS>(N, i1, i2, segs + n, segs + n + N, N);
      if (ns == 0) return -1;
      for (int k = 0; k < ns; k++) segs[n + ns + k] = segs[n + N + k];  // w after o
    }
    D->segoff[d] = n;
    D->nseg[d] = ns;
    n += 2 * ns;
  }
  return n;
}

// pencil groups of any sweep launch: ceil(n_t1 / kTP) * n_t2
long long fallback_groups(const int GD[3]) {
  long long ngrp = 1;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (a != b) ngrp = std::max(ngrp, (long long)((GD[a] + kTP - 1) / kTP) * GD[b]);
  return ngrp;
}

int launch_pressure(hipPPMData *D, const hipPPMParameter *P, double min_support, int want_dt,
                    double cfl) {
  PresParams Q;
  memset(&Q, 0, sizeof(Q));
  Q.rank = (int)P->GridRank;
  Q.idual = (int)P->DualEnergyFormalism;
  for (int d = 0; d < 3; d++) {
    Q.GD[d] = D->GD[d];
    Q.st[d] = (int)P->GridStartIndex[d];
    Q.en[d] = (int)P->GridEndIndex[d];
    Q.cw[d] = D->comoving ? D->cwdt[d] : D->cwdev[d];
  }
  Q.cells = D->cells;
  Q.A = D->A;
  Q.pres = D->pres;
  Q.gamma = P->Gamma;
  Q.tiny_number = P->MinimumPressure;
  Q.min_support = min_support;
  Q.aye = D->comoving ? D->aye : 1.0;
  Q.want_dt = want_dt;
  Q.dtmin = D->dtmin;
  if (want_dt) {
    hipLaunchKernelGGL(dt_init_kernel, dim3(1), dim3(64), 0, D->stream, D->dtmin);
    HIPCHK(hipGetLastError());
  }
  // at most 8192 workgroups (a grid-stride loop: at 256^3, 188 us per pass
  // with the Courant step against 204-218 us with one cell per work-item)
  long long nb = (D->cells + 255) / 256;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(pressure_dt_kernel, dim3((unsigned)nb), dim3(256), 0, D->stream, Q);
  HIPCHK(hipGetLastError());
  if (want_dt) {
    hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(64), 0, D->stream, D->dtmin, cfl, D->dt);
    HIPCHK(hipGetLastError());
  }
  return ENZO_PPM_SUCCESS;
}


}  // namespace

// ==========================================================================
extern "C" {

int hipPPMInitParameter(hipPPMParameter *Para, enzo_int DensNum, enzo_int TENum,
                        enzo_int Vel1Num, enzo_int Vel2Num, enzo_int Vel3Num, enzo_int GENum,
                        enzo_int MaxNumberOfBaryonFields, const enzo_long_int GridGlobalStart[],
                        const enzo_int GridStartIndex[], const enzo_int GridEndIndex[],
                        const enzo_int GridDimension[], enzo_float dxx, enzo_float dxy,
                        enzo_float dxz, enzo_float Gamma, enzo_int PPMDiffusionParameter,
                        enzo_int PPMFlatteningParameter, enzo_int PPMSteepeningParameter,
                        enzo_int ConservativeReconstruction, enzo_int PositiveReconstruction,
                        enzo_int GravityOn, enzo_int DualEnergyFormalism,
                        enzo_float DualEnergyFormalismEta1, enzo_float DualEnergyFormalismEta2,
                        enzo_int PressureFree, enzo_float MinimumPressure,
                        enzo_int RiemannSolver, enzo_int RiemannSolverFallback,
                        enzo_int NumberOfColours, const enzo_int *colnum, enzo_int GridRank,
                        enzo_float SmallRho) {
  if (!Para || !GridStartIndex || !GridEndIndex || !GridDimension) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfColours < 0 || NumberOfColours > ENZO_PPM_MAX_COLOR) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfColours > 0 && !colnum) return ENZO_PPM_ERR_ARGUMENT;
  memset(Para, 0, sizeof(*Para));
  Para->DensNum = DensNum; Para->TENum = TENum; Para->Vel1Num = Vel1Num;
  Para->Vel2Num = Vel2Num; Para->Vel3Num = Vel3Num; Para->GENum = GENum;
  Para->MaxNumberOfBaryonFields = MaxNumberOfBaryonFields;
  for (int d = 0; d < 3; d++) {
    Para->GridGlobalStart[d] = GridGlobalStart ? GridGlobalStart[d] : 0;
    Para->GridStartIndex[d] = GridStartIndex[d];
    Para->GridEndIndex[d] = GridEndIndex[d];
    Para->GridDimension[d] = GridDimension[d];
  }
  Para->GridRank = GridRank;
  Para->dx[0] = dxx; Para->dx[1] = dxy; Para->dx[2] = dxz;
  Para->Gamma = Gamma;
  Para->PPMDiffusionParameter = PPMDiffusionParameter;
  Para->PPMFlatteningParameter = PPMFlatteningParameter;
  Para->PPMSteepeningParameter = PPMSteepeningParameter;
  Para->ConservativeReconstruction = ConservativeReconstruction;
  Para->PositiveReconstruction = PositiveReconstruction;
  Para->GravityOn = GravityOn;
  Para->DualEnergyFormalism = DualEnergyFormalism;
  Para->DualEnergyFormalismEta1 = DualEnergyFormalismEta1;
  Para->DualEnergyFormalismEta2 = DualEnergyFormalismEta2;
  Para->PressureFree = PressureFree;
  Para->MinimumPressure = MinimumPressure;
  Para->RiemannSolver = RiemannSolver;
  Para->RiemannSolverFallback = RiemannSolverFallback;
  Para->SmallRho = SmallRho;
  Para->NumberOfColours = NumberOfColours;
  for (int n = 0; n < NumberOfColours; n++) Para->colnum[n] = colnum[n];
  return check_para(Para);
}

int hipPPMSetCellWidth(hipPPMParameter *Para, enzo_int dim, const enzo_float *CellWidthTemp) {
  if (!Para || dim < 0 || dim > 2) return ENZO_PPM_ERR_ARGUMENT;
  Para->CellWidth[dim] = CellWidthTemp;
  return ENZO_PPM_SUCCESS;
}

int hipPPMInitData(hipPPMData **Data, const hipPPMParameter *Para, void *stream) {
  if (!Data) return ENZO_PPM_ERR_ARGUMENT;
  int st = check_para(Para);
  if (st != ENZO_PPM_SUCCESS) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ENZO_PPM_ERR_NO_DEVICE;
  hipPPMData *D = *Data;
  long long cells = 1;
  for (int d = 0; d < 3; d++) cells *= Para->GridDimension[d];
  const int nslot = nslot_of(Para);
  const bool same = D && D->cells == cells && D->nslot == nslot &&
                    D->GD[0] == Para->GridDimension[0] && D->GD[1] == Para->GridDimension[1] &&
                    D->GD[2] == Para->GridDimension[2] && (D->gravity || !Para->GravityOn);
  if (!same) {
    if (D) {
      free_data(D);
      delete D;
    }
    D = new (std::nothrow) hipPPMData();
    if (!D) return ENZO_PPM_ERR_ARGUMENT;
    *Data = D;
    HIPCHK(hipGetDevice(&D->device));
    for (int d = 0; d < 3; d++) D->GD[d] = (int)Para->GridDimension[d];
    D->cells = cells;
    D->nslot = nslot;
    D->ncol = (int)Para->NumberOfColours;
    D->gravity = Para->GravityOn != 0;
    size_t fb = sizeof(double) * (size_t)cells;
    if (alloc_dev((void **)&D->A, fb * nslot) || alloc_dev((void **)&D->B, fb * nslot) ||
        alloc_dev((void **)&D->pres, fb) ||
        (D->gravity && alloc_dev((void **)&D->accel, 3 * fb)))
      return ENZO_PPM_ERR_HIP;
    long long fcap = 0;
    for (int d = 0; d < 3; d++) {
      if (alloc_dev((void **)&D->tab[d], sizeof(double) * NTAB * D->GD[d]) ||
          alloc_dev((void **)&D->cwdev[d], sizeof(double) * D->GD[d]))
        return ENZO_PPM_ERR_HIP;
      fcap = std::max(fcap, (long long)D->GD[d] * D->GD[(d + 2) % 3]);
    }
    D->flat_cap = fcap;
    int nmax = std::max(D->GD[0], std::max(D->GD[1], D->GD[2]));
    if (alloc_dev((void **)&D->flat, sizeof(double) * fcap) ||
        alloc_dev((void **)&D->qc, sizeof(double) * nmax) ||
        alloc_dev((void **)&D->dt, sizeof(double)) ||
        alloc_dev((void **)&D->dtmin, sizeof(unsigned long long)) ||
        alloc_dev((void **)&D->counters, 4 * sizeof(unsigned long long)))
      return ENZO_PPM_ERR_HIP;
    const long long ngrp = fallback_groups(D->GD);
    D->fb_groups = ngrp;
    if (alloc_dev((void **)&D->fb_flag, sizeof(int) * ngrp) ||
        alloc_dev((void **)&D->fb_list, sizeof(unsigned) * ngrp) ||
        alloc_dev((void **)&D->fb_state, 2 * sizeof(unsigned)))
      return ENZO_PPM_ERR_HIP;
  }
  D->stream = (hipStream_t)stream;
  // geometry tables and segments (cheap; rebuilt on every init)
  st = upload_geometry(D, Para);
  if (st != ENZO_PPM_SUCCESS) return st;
  std::vector<int> segs(2 * 3 * (size_t)ENZO_PPM_MAX_ANY_SINGLE_DIRECTION);
  const int nsegs = plan_segments(D, Para, segs.data());
  if (nsegs < 0) return ENZO_PPM_ERR_ARGUMENT;
  if (D->segbuf) (void)hipFree(D->segbuf);
  D->segbuf = nullptr;
  if (alloc_dev((void **)&D->segbuf, sizeof(int) * std::max(1, nsegs)))
    return ENZO_PPM_ERR_HIP;
  if (nsegs > 0)
    HIPCHK(hipMemcpy(D->segbuf, segs.data(), sizeof(int) * nsegs, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, D->stream, D->counters, 4);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(D->fb_flag, 0, sizeof(int) * D->fb_groups, D->stream));
  HIPCHK(hipMemsetAsync(D->fb_state, 0, 2 * sizeof(unsigned), D->stream));
  D->fallback_total = 0;
  D->chain_total = 0;
  return ENZO_PPM_SUCCESS;
}

int hipPPMSetBaryon(hipPPMData *D, const hipPPMParameter *P, enzo_float *BaryonField[],
                    enzo_float *Pressure, enzo_float *AccelerationField[]) {
  if (!D || !P || !BaryonField) return ENZO_PPM_ERR_ARGUMENT;
  std::vector<long long> f;
  slot_fields(P, f);
  const size_t fb = sizeof(double) * (size_t)D->cells;
  std::vector<HostXfer> up;
  for (int s = 0; s < D->nslot; s++) {
    double *dst = D->A + (size_t)s * D->cells;
    if (f[s] >= 0) {
      if (!BaryonField[f[s]]) return ENZO_PPM_ERR_ARGUMENT;
      up.push_back({(char *)dst, (const char *)BaryonField[f[s]], fb});
    } else {
      HIPCHK(hipMemsetAsync(dst, 0, fb, D->stream));
    }
  }
  if (P->GravityOn) {
    if (!AccelerationField || !D->accel) return ENZO_PPM_ERR_ARGUMENT;
    for (int d = 0; d < 3; d++) {
      if (d < P->GridRank && AccelerationField[d])
        up.push_back({(char *)(D->accel + (size_t)d * D->cells), (const char *)AccelerationField[d], fb});
      else
        HIPCHK(hipMemsetAsync(D->accel + (size_t)d * D->cells, 0, fb, D->stream));
    }
  }
  if (Pressure) up.push_back({(char *)D->pres, (const char *)Pressure, fb});
  const int sup = stage_upload(D, up);
  if (sup != ENZO_PPM_SUCCESS) return sup;
  if (!Pressure) {
    int st = launch_pressure(D, P, 0.0, 0, 1.0);
    if (st != ENZO_PPM_SUCCESS) return st;
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMSetSubgridFluxes(hipPPMData *D, const hipPPMParameter *P, enzo_fluxes *SubgridFluxes[],
                           enzo_int NumberOfSubgrids, const enzo_long_int GridGlobalStart[]) {
  if (!D || !P || NumberOfSubgrids < 0) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfSubgrids > 0 && (!SubgridFluxes || !GridGlobalStart)) return ENZO_PPM_ERR_ARGUMENT;
  D->nsub = (int)NumberOfSubgrids;
  const int nfld = nslot_of(P);
  for (int dim = 0; dim < 3; dim++) {
    D->hfaces[dim].clear();
    D->hface_sub[dim].clear();
    D->hface_side[dim].clear();
    if (D->face_head[dim]) (void)hipFree(D->face_head[dim]);
    if (D->faces[dim]) (void)hipFree(D->faces[dim]);
    if (D->face_out[dim]) (void)hipFree(D->face_out[dim]);
    D->face_head[dim] = nullptr;
    D->faces[dim] = nullptr;
    D->face_out[dim] = nullptr;
    D->face_total[dim] = 0;
    std::vector<int> head;
    long long total = 0;
    const int st = build_face_records(D->GD, P, SubgridFluxes, (int)NumberOfSubgrids,
                                      GridGlobalStart, nfld, dim, D->hfaces[dim],
                                      D->hface_sub[dim], D->hface_side[dim], head, total);
    if (st != ENZO_PPM_SUCCESS) return st;
    const std::vector<FaceRec> &sorted = D->hfaces[dim];
    D->face_total[dim] = total;
    if (total > 0) {
      if (alloc_dev((void **)&D->face_head[dim], sizeof(int) * head.size()) ||
          alloc_dev((void **)&D->faces[dim], sizeof(FaceRec) * sorted.size()) ||
          alloc_dev((void **)&D->face_out[dim], sizeof(double) * total))
        return ENZO_PPM_ERR_HIP;
      HIPCHK(hipMemcpy(D->face_head[dim], head.data(), sizeof(int) * head.size(),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(D->faces[dim], sorted.data(), sizeof(FaceRec) * sorted.size(),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemsetAsync(D->face_out[dim], 0, sizeof(double) * total, D->stream));
    }
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMSweep(hipPPMData *D, const hipPPMParameter *P, enzo_float dtFixed, enzo_int dir) {
  if (!D || !P || dir < 0 || dir > 2 || dir >= P->GridRank) return ENZO_PPM_ERR_ARGUMENT;
  if (D->nseg[dir] == 0) return ENZO_PPM_ERR_ARGUMENT;
  return launch_sweep(D, P, dtFixed, nullptr, (int)dir);
}

int hipPPMSaveSubgridFluxes(enzo_fluxes *SubgridFluxes[], enzo_int NumberOfSubgrids,
                            const enzo_long_int GridGlobalStart[], hipPPMData *D,
                            const hipPPMParameter *P, enzo_int dir) {
  (void)GridGlobalStart;
  if (!D || !P || dir < 0 || dir > 2) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfSubgrids != D->nsub) return ENZO_PPM_ERR_ARGUMENT;
  if (D->face_total[dir] == 0) return ENZO_PPM_SUCCESS;
  std::vector<double> h((size_t)D->face_total[dir]);
  HIPCHK(hipMemcpyAsync(h.data(), D->face_out[dir], sizeof(double) * h.size(),
                        hipMemcpyDeviceToHost, D->stream));
  HIPCHK(hipStreamSynchronize(D->stream));
  const long long nact[3] = {P->GridEndIndex[0] - P->GridStartIndex[0] + 1,
                             P->GridEndIndex[1] - P->GridStartIndex[1] + 1,
                             P->GridEndIndex[2] - P->GridStartIndex[2] + 1};
  std::vector<long long> f;
  slot_fields(P, f);
  scatter_face_fluxes(h, D->hfaces[dir], D->hface_sub[dir], D->hface_side[dir], f, nact, (int)dir,
                      SubgridFluxes);
  return ENZO_PPM_SUCCESS;
}

int hipPPMGetBaryon(hipPPMData *D, const hipPPMParameter *P, enzo_float *BaryonField[]) {
  if (!D || !P || !BaryonField) return ENZO_PPM_ERR_ARGUMENT;
  std::vector<long long> f;
  slot_fields(P, f);
  const size_t fb = sizeof(double) * (size_t)D->cells;
  std::vector<HostXfer> down;
  for (int s = 0; s < D->nslot; s++) {
    if (f[s] < 0 || !BaryonField[f[s]]) continue;
    down.push_back({(char *)BaryonField[f[s]], (const char *)(D->A + (size_t)s * D->cells), fb});
  }
  const int sdn = stage_download(D, down);
  if (sdn != ENZO_PPM_SUCCESS) return sdn;
  unsigned long long c[4];
  HIPCHK(hipMemcpy(c, D->counters, sizeof(c), hipMemcpyDeviceToHost));
  D->fallback_total = (long long)c[0];
  D->chain_total = (long long)c[1];
  if (c[3] != 0) {
    snprintf(g_last_error, sizeof(g_last_error),
             "the per-cell fallback was needed on a pencil longer than %d cells", kFbTS + 1);
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  }
  if (c[2] != 0) return ENZO_PPM_ERR_NEGATIVE;
  if (c[0] != 0 && P->RiemannSolverFallback == 0) return ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}

int hipPPMDestroy(hipPPMData *D) {
  if (!D) return ENZO_PPM_SUCCESS;
  free_data(D);
  delete D;
  return ENZO_PPM_SUCCESS;
}

enzo_long_int hipPPMFallbackCount(hipPPMData *D) {
  if (!D) return -1;
  unsigned long long c[4];
  if (hipMemcpy(c, D->counters, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return (enzo_long_int)c[0];
}

int hipPPMSolvePPM_DE(const enzo_ppm_grid *g, const enzo_ppm_globals *gl, enzo_int CycleNumber,
                      enzo_int NumberOfSubgrids, enzo_fluxes *SubgridFluxes[],
                      enzo_float *CellWidthTemp[], const enzo_long_int GridGlobalStart[],
                      enzo_int GravityOn, enzo_int NumberOfColours, const enzo_int colnum[],
                      enzo_float MinimumSupportEnergyCoefficient, hipPPMData **Data,
                      void *stream) {
  if (!g || !gl) return ENZO_PPM_ERR_ARGUMENT;
  if (gl->EOSType != 0 || gl->ReconstructionMethod != ENZO_PPM_RECONSTRUCTION_PPM)
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  hipPPMParameter Para;
  enzo_long_int ggs[3] = {0, 0, 0};
  if (GridGlobalStart)
    for (int d = 0; d < 3; d++) ggs[d] = GridGlobalStart[d];
  int st = hipPPMInitParameter(
      &Para, g->DensNum, g->TENum, g->Vel1Num, g->Vel2Num, g->Vel3Num, g->GENum,
      ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS, ggs, g->GridStartIndex, g->GridEndIndex,
      g->GridDimension, CellWidthTemp ? CellWidthTemp[0][0] : 1.0,
      CellWidthTemp ? CellWidthTemp[1][0] : 1.0, CellWidthTemp ? CellWidthTemp[2][0] : 1.0,
      gl->Gamma, gl->PPMDiffusionParameter, gl->PPMFlatteningParameter,
      gl->PPMSteepeningParameter, gl->ConservativeReconstruction, gl->PositiveReconstruction,
      GravityOn, gl->DualEnergyFormalism, gl->DualEnergyFormalismEta1,
      gl->DualEnergyFormalismEta2, gl->PressureFree, gl->tiny_number, gl->RiemannSolver,
      gl->RiemannSolverFallback, NumberOfColours, colnum, g->GridRank, gl->SmallRho);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (CellWidthTemp)
    for (int d = 0; d < 3; d++) hipPPMSetCellWidth(&Para, d, CellWidthTemp[d]);
  hipPPMData *local = nullptr;
  hipPPMData **pd = Data ? Data : &local;
  st = hipPPMInitData(pd, &Para, stream);
  if (st != ENZO_PPM_SUCCESS) return st;
  hipPPMData *D = *pd;
  // stale pressure (Grid_SolvePPM_DE.C:58-61)
  st = hipPPMSetBaryon(D, &Para, g->BaryonField, nullptr, g->AccelerationField);
  if (st == ENZO_PPM_SUCCESS && MinimumSupportEnergyCoefficient != 0.0)
    st = launch_pressure(D, &Para, MinimumSupportEnergyCoefficient, 0, 1.0);
  if (st == ENZO_PPM_SUCCESS)
    st = hipPPMSetSubgridFluxes(D, &Para, SubgridFluxes, NumberOfSubgrids, ggs);
  const long long nact[3] = {g->GridEndIndex[0] - g->GridStartIndex[0] + 1,
                             g->GridEndIndex[1] - g->GridStartIndex[1] + 1,
                             g->GridEndIndex[2] - g->GridStartIndex[2] + 1};
  const int rank = (int)g->GridRank;
  const int ixyz = (int)(CycleNumber % rank);
  for (int n = ixyz; n < ixyz + rank && st == ENZO_PPM_SUCCESS; n++) {
    const int dir = n % rank;
    if (nact[dir] <= 1) continue;
    st = launch_sweep(D, &Para, g->dtFixed, nullptr, dir);
    if (st == ENZO_PPM_SUCCESS && NumberOfSubgrids > 0)
      st = hipPPMSaveSubgridFluxes(SubgridFluxes, NumberOfSubgrids, ggs, D, &Para, dir);
  }
  if (st == ENZO_PPM_SUCCESS) st = hipPPMGetBaryon(D, &Para, g->BaryonField);
  if (!Data) hipPPMDestroy(local);
  return st;
}

}  // extern "C"

// ------------------------------------------------------------------------
// The entry point that allocates nothing (hipPPMSolvePPM_DE_NoAlloc): its
// device buffers are carved from `scratch`, one static array per GPU that
// the GPU holds until the process ends.  Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current.
constexpr size_t kScratchBytes = size_t(ENZO_PPM_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the current device's scratch address (looked up on its first use)
int scratch_of_current_device(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ENZO_PPM_ERR_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ENZO_PPM_SUCCESS;
}

// consecutive 256-byte aligned pieces of the scratch array
struct ScratchCarve {
  unsigned char *base;
  size_t used = 0;
  template <class T>
  T *take(size_t n) {
    used = (used + 255) & ~(size_t)255;
    T *p = reinterpret_cast<T *>(base + used);
    used += n * sizeof(T);
    return p;
  }
};

unsigned grid_blocks(long long items) {
  long long nb = (items + 255) / 256;
  return (unsigned)std::max(1LL, std::min(nb, 8192LL));
}

}  // namespace

extern "C" {

int hipPPMSolvePPM_DE_NoAlloc(const enzo_ppm_grid *g, const enzo_ppm_globals *gl,
                              enzo_int CycleNumber, enzo_float *CellWidthTemp[],
                              enzo_int GravityOn, enzo_int NumberOfColours, const enzo_int colnum[],
                              enzo_float MinimumSupportEnergyCoefficient, void *stream,
                              enzo_long_int *FallbackCells) {
  if (!g || !gl || !g->BaryonField) return ENZO_PPM_ERR_ARGUMENT;
  if (gl->EOSType != 0 || gl->ReconstructionMethod != ENZO_PPM_RECONSTRUCTION_PPM)
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ENZO_PPM_ERR_NO_DEVICE;
  hipPPMParameter Para;
  const enzo_long_int ggs[3] = {0, 0, 0};
  int st = hipPPMInitParameter(
      &Para, g->DensNum, g->TENum, g->Vel1Num, g->Vel2Num, g->Vel3Num, g->GENum,
      ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS, ggs, g->GridStartIndex, g->GridEndIndex,
      g->GridDimension, CellWidthTemp ? CellWidthTemp[0][0] : 1.0,
      CellWidthTemp ? CellWidthTemp[1][0] : 1.0, CellWidthTemp ? CellWidthTemp[2][0] : 1.0,
      gl->Gamma, gl->PPMDiffusionParameter, gl->PPMFlatteningParameter,
      gl->PPMSteepeningParameter, gl->ConservativeReconstruction, gl->PositiveReconstruction,
      GravityOn, gl->DualEnergyFormalism, gl->DualEnergyFormalismEta1,
      gl->DualEnergyFormalismEta2, gl->PressureFree, gl->tiny_number, gl->RiemannSolver,
      gl->RiemannSolverFallback, NumberOfColours, colnum, g->GridRank, gl->SmallRho);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (CellWidthTemp)
    for (int d = 0; d < 3; d++) hipPPMSetCellWidth(&Para, d, CellWidthTemp[d]);
  st = check_para(&Para);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (GravityOn && !g->AccelerationField) return ENZO_PPM_ERR_ARGUMENT;

  hipPPMData D;  // no device buffers of its own: every pointer is into scratch
  for (int d = 0; d < 3; d++) D.GD[d] = (int)Para.GridDimension[d];
  D.cells = (long long)D.GD[0] * D.GD[1] * D.GD[2];
  D.nslot = nslot_of(&Para);
  D.ncol = (int)Para.NumberOfColours;
  D.gravity = GravityOn != 0;
  D.stream = (hipStream_t)stream;
  HIPCHK(hipGetDevice(&D.device));
  void *base = nullptr;
  st = scratch_of_current_device(&base);
  if (st != ENZO_PPM_SUCCESS) return st;
  ScratchCarve sc{static_cast<unsigned char *>(base)};
  const size_t cells = (size_t)D.cells;
  D.A = sc.take<double>(cells * D.nslot);
  D.B = sc.take<double>(cells * D.nslot);
  D.pres = sc.take<double>(cells);
  long long fcap = 0;
  for (int d = 0; d < 3; d++) {
    D.tab[d] = sc.take<double>((size_t)NTAB * D.GD[d]);
    D.cwdev[d] = sc.take<double>((size_t)D.GD[d]);
    fcap = std::max(fcap, (long long)D.GD[d] * D.GD[(d + 2) % 3]);
  }
  D.flat_cap = fcap;
  D.flat = sc.take<double>((size_t)fcap);
  D.qc = sc.take<double>((size_t)std::max(D.GD[0], std::max(D.GD[1], D.GD[2])));
  D.dt = sc.take<double>(1);
  D.dtmin = sc.take<unsigned long long>(1);
  D.counters = sc.take<unsigned long long>(4);
  D.fb_groups = fallback_groups(D.GD);
  D.fb_flag = sc.take<int>((size_t)D.fb_groups);
  D.fb_list = sc.take<unsigned>((size_t)D.fb_groups);
  D.fb_state = sc.take<unsigned>(2);
  static thread_local int segs[2 * 3 * ENZO_PPM_MAX_ANY_SINGLE_DIRECTION];
  const int nsegs = plan_segments(&D, &Para, segs);
  if (nsegs < 0) return ENZO_PPM_ERR_ARGUMENT;
  D.segbuf = sc.take<int>((size_t)std::max(1, nsegs));
  if (sc.used > kScratchBytes) {
    snprintf(g_last_error, sizeof(g_last_error),
             "hipPPMSolvePPM_DE_NoAlloc: the grid needs %zu bytes of scratch, more than its "
             "%zu (%d MiB)", sc.used, kScratchBytes, ENZO_PPM_NOALLOC_SCRATCH_MIB);
    return ENZO_PPM_ERR_ARGUMENT;
  }
  if (GravityOn)
    for (int d = 0; d < 3; d++) {
      if (d < Para.GridRank && !g->AccelerationField[d]) return ENZO_PPM_ERR_ARGUMENT;
      D.accel_ext[d] = d < Para.GridRank ? g->AccelerationField[d] : nullptr;
    }
  SlotPtrs F;
  for (int s = 0; s < D.nslot; s++) {
    const long long f = slot_field(&Para, s);
    F.f[s] = f >= 0 ? g->BaryonField[f] : nullptr;
    if (f >= 0 && !F.f[s]) return ENZO_PPM_ERR_ARGUMENT;
  }
  // scratch keeps the last call's contents: clear what must start at zero
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, D.stream, D.counters, 4);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_blocks(D.fb_groups)), dim3(256), 0, D.stream,
                     reinterpret_cast<unsigned *>(D.fb_flag), D.fb_groups);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, D.stream, D.fb_state, 2LL);
  HIPCHK(hipGetLastError());
  st = upload_geometry(&D, &Para);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (nsegs > 0) {
    HIPCHK(hipMemcpyAsync(D.segbuf, segs, sizeof(int) * nsegs, hipMemcpyHostToDevice, D.stream));
    HIPCHK(hipStreamSynchronize(D.stream));  // segs is reused by the next call of this thread
  }
  hipLaunchKernelGGL(gather_slots_kernel, dim3(grid_blocks((long long)cells * D.nslot)), dim3(256),
                     0, D.stream, D.A, D.cells, D.nslot, F);
  HIPCHK(hipGetLastError());
  // stale pressure (Grid_SolvePPM_DE.C:58-61), as hipPPMSolvePPM_DE takes it
  st = launch_pressure(&D, &Para, 0.0, 0, 1.0);
  if (st == ENZO_PPM_SUCCESS && MinimumSupportEnergyCoefficient != 0.0)
    st = launch_pressure(&D, &Para, MinimumSupportEnergyCoefficient, 0, 1.0);
  const int rank = (int)g->GridRank;
  const int ixyz = (int)(CycleNumber % rank);
  for (int n = ixyz; n < ixyz + rank && st == ENZO_PPM_SUCCESS; n++) {
    const int dir = n % rank;
    if (g->GridEndIndex[dir] - g->GridStartIndex[dir] + 1 <= 1) continue;
    st = launch_sweep(&D, &Para, g->dtFixed, nullptr, dir);
  }
  if (st != ENZO_PPM_SUCCESS) return st;
  hipLaunchKernelGGL(scatter_slots_kernel, dim3(grid_blocks((long long)cells * D.nslot)),
                     dim3(256), 0, D.stream, D.A, D.cells, D.nslot, F);
  HIPCHK(hipGetLastError());
  unsigned long long c[4];
  HIPCHK(hipMemcpyAsync(c, D.counters, sizeof(c), hipMemcpyDeviceToHost, D.stream));
  HIPCHK(hipStreamSynchronize(D.stream));
  if (FallbackCells) *FallbackCells = (enzo_long_int)c[0];
  if (c[3] != 0) {
    snprintf(g_last_error, sizeof(g_last_error),
             "the per-cell fallback was needed on a pencil longer than %d cells", kFbTS + 1);
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  }
  if (c[2] != 0) return ENZO_PPM_ERR_NEGATIVE;
  if (c[0] != 0 && Para.RiemannSolverFallback == 0) return ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}

enzo_float *hipPPMDeviceField(hipPPMData *D, enzo_int field) {
  if (!D) return nullptr;
  if (field == -1) return D->pres;
  if (field < 0 || field >= D->nslot) return nullptr;
  return D->A + (size_t)field * D->cells;
}

int hipPPMPressureAndTimeStep(hipPPMData *D, const hipPPMParameter *P,
                              enzo_float CourantSafetyNumber,
                              enzo_float MinimumSupportEnergyCoefficient, enzo_float *dt_host) {
  if (!D || !P) return ENZO_PPM_ERR_ARGUMENT;
  int st = launch_pressure(D, P, MinimumSupportEnergyCoefficient, 1, CourantSafetyNumber);
  if (st != ENZO_PPM_SUCCESS) return st;
  if (dt_host) {
    HIPCHK(hipMemcpyAsync(dt_host, D->dt, sizeof(double), hipMemcpyDeviceToHost, D->stream));
    HIPCHK(hipStreamSynchronize(D->stream));
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMStep(hipPPMData *D, const hipPPMParameter *P, enzo_int CycleNumber,
               enzo_float dtFixed) {
  if (!D || !P) return ENZO_PPM_ERR_ARGUMENT;
  const int rank = (int)P->GridRank;
  if (D->comoving) {  // OldBaryonField of the time-centred expansion pressure
    int slots[5];
    const int ns = expansion_old_slots((int)P->DualEnergyFormalism, rank, slots);
    for (int k = 0; k < ns; k++)
      HIPCHK(hipMemcpyAsync(D->old + (size_t)slots[k] * D->cells, D->A + (size_t)slots[k] * D->cells,
                            sizeof(double) * (size_t)D->cells, hipMemcpyDeviceToDevice, D->stream));
    D->old_valid = true;
  }
  const int ixyz = (int)(CycleNumber % rank);
  for (int n = ixyz; n < ixyz + rank; n++) {
    const int dir = n % rank;
    if (P->GridEndIndex[dir] - P->GridStartIndex[dir] + 1 <= 1) continue;
    int st = launch_sweep(D, P, dtFixed, dtFixed > 0 ? nullptr : D->dt, dir);
    if (st != ENZO_PPM_SUCCESS) return st;
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMSetComoving(hipPPMData *D, const hipPPMParameter *P, enzo_int enable) {
  if (!D || !P) return ENZO_PPM_ERR_ARGUMENT;
  HIPCHK(hipStreamSynchronize(D->stream));
  if (enable) {
    for (int d = 0; d < 3; d++)
      if (!D->cwdt[d] && alloc_dev((void **)&D->cwdt[d], sizeof(double) * D->GD[d]))
        return ENZO_PPM_ERR_HIP;
    if (!D->old && alloc_dev((void **)&D->old, sizeof(double) * (size_t)(S_GE + 1) * D->cells))
      return ENZO_PPM_ERR_HIP;
  }
  D->comoving = enable != 0;
  D->old_valid = false;
  if (!D->comoving) D->a_sweep = D->aye = 1.0;
  return upload_geometry(D, P);
}

int hipPPMSetExpansionFactors(hipPPMData *D, const hipPPMParameter *P, enzo_float a_courant,
                              enzo_float a_sweep) {
  if (!D || !P || !D->comoving || !(a_courant > 0.0) || !(a_sweep > 0.0))
    return ENZO_PPM_ERR_ARGUMENT;
  D->aye = a_courant;
  if (a_sweep == D->a_sweep) return ENZO_PPM_SUCCESS;
  D->a_sweep = a_sweep;
  return upload_geometry(D, P);
}

int hipPPMComovingExpansionTerms(hipPPMData *D, const hipPPMParameter *P, enzo_float coef,
                                 enzo_float Coefficient) {
  if (!D || !P || !D->comoving) return ENZO_PPM_ERR_ARGUMENT;
  ExpandParams X;
  X.cells = D->cells;
  X.rank = (int)P->GridRank;
  X.idual = (int)P->DualEnergyFormalism;
  X.interp = D->old_valid && coef != 1.0 ? 1 : 0;
  X.gamma = P->Gamma;
  X.tiny_number = P->MinimumPressure;
  X.coef = coef;
  X.Coefficient = Coefficient;
  X.A = D->A;
  X.old = D->old;
  // one cell per work-item: 2.9 ms per 518^3 pass against 3.4 ms with a
  // grid-stride loop over 16384 workgroups
  const long long nb = (D->cells + 255) / 256;
  hipLaunchKernelGGL(expand_kernel, dim3((unsigned)nb), dim3(256), 0, D->stream, X);
  HIPCHK(hipGetLastError());
  return ENZO_PPM_SUCCESS;
}

int hipPPMProjectBoundaryFluxes(const enzo_fluxes *Fine, enzo_int rank, enzo_int nfields,
                                const enzo_int *Refinement, enzo_fluxes *Projected) {
  if (!Fine || !Projected || !Refinement || rank < 1 || rank > 3 || nfields < 1 ||
      nfields > ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS)
    return ENZO_PPM_ERR_ARGUMENT;
  for (int i = 0; i < 3; i++)
    if (Refinement[i] < 1) return ENZO_PPM_ERR_ARGUMENT;
  for (int dim = 0; dim < rank; dim++)
    for (int i = 0; i < rank; i++) {
      Projected->LeftFluxStartGlobalIndex[dim][i] = Fine->LeftFluxStartGlobalIndex[dim][i] / Refinement[i];
      Projected->LeftFluxEndGlobalIndex[dim][i] = Fine->LeftFluxEndGlobalIndex[dim][i] / Refinement[i];
      Projected->RightFluxStartGlobalIndex[dim][i] = Fine->RightFluxStartGlobalIndex[dim][i] / Refinement[i];
      Projected->RightFluxEndGlobalIndex[dim][i] = Fine->RightFluxEndGlobalIndex[dim][i] / Refinement[i];
    }
  int st = ENZO_PPM_SUCCESS;
  for (int dim = 0; dim < rank && st == ENZO_PPM_SUCCESS; dim++) {
    ProjParams Q;
    long long TotalRefinement = 1;
    for (int i = 0; i < 3; i++) {
      Q.Dims[i] = i < rank ? (int)(Fine->LeftFluxEndGlobalIndex[dim][i] -
                                   Fine->LeftFluxStartGlobalIndex[dim][i] + 1) : 1;
      Q.r[i] = (int)Refinement[i];
      TotalRefinement *= Refinement[i];
    }
    for (int i = 0; i < 3; i++) {
      Q.PD[i] = i != dim ? Q.Dims[i] / Q.r[i] : 1;
      if (i != dim && i < rank && Q.Dims[i] % Q.r[i] != 0) st = ENZO_PPM_ERR_ARGUMENT;
    }
    if (st != ENZO_PPM_SUCCESS) break;
    Q.nfine = (long long)Q.Dims[0] * Q.Dims[1] * Q.Dims[2];
    Q.ncoarse = (long long)Q.PD[0] * Q.PD[1] * Q.PD[2];
    Q.narrays = 2 * (int)nfields;
    Q.dArea = 1.0 / double(TotalRefinement);
    std::vector<double> fine((size_t)Q.narrays * Q.nfine), coarse((size_t)Q.narrays * Q.ncoarse);
    for (int f = 0; f < nfields && st == ENZO_PPM_SUCCESS; f++) {
      const enzo_float *L = Fine->LeftFluxes[f][dim], *R = Fine->RightFluxes[f][dim];
      if (!L || !R || !Projected->LeftFluxes[f][dim] || !Projected->RightFluxes[f][dim]) {
        st = ENZO_PPM_ERR_ARGUMENT;
        break;
      }
      std::copy(L, L + Q.nfine, fine.begin() + (size_t)(2 * f) * Q.nfine);
      std::copy(R, R + Q.nfine, fine.begin() + (size_t)(2 * f + 1) * Q.nfine);
    }
    if (st != ENZO_PPM_SUCCESS) break;
    double *dfine = nullptr, *dcoarse = nullptr;
    if (alloc_dev((void **)&dfine, sizeof(double) * fine.size()) ||
        alloc_dev((void **)&dcoarse, sizeof(double) * coarse.size())) {
      st = ENZO_PPM_ERR_HIP;
    } else if (hipMemcpy(dfine, fine.data(), sizeof(double) * fine.size(), hipMemcpyHostToDevice) !=
               hipSuccess) {
      st = ENZO_PPM_ERR_HIP;
    } else {
      Q.fine = dfine;
      Q.coarse = dcoarse;
      const long long n = (long long)Q.narrays * Q.ncoarse;
      hipLaunchKernelGGL(project_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, Q);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpy(coarse.data(), dcoarse, sizeof(double) * coarse.size(), hipMemcpyDeviceToHost) !=
              hipSuccess)
        st = ENZO_PPM_ERR_HIP;
    }
    if (dfine) (void)hipFree(dfine);
    if (dcoarse) (void)hipFree(dcoarse);
    if (st != ENZO_PPM_SUCCESS) break;
    for (int f = 0; f < nfields; f++) {
      std::copy(coarse.begin() + (size_t)(2 * f) * Q.ncoarse,
                coarse.begin() + (size_t)(2 * f + 1) * Q.ncoarse, Projected->LeftFluxes[f][dim]);
      std::copy(coarse.begin() + (size_t)(2 * f + 1) * Q.ncoarse,
                coarse.begin() + (size_t)(2 * f + 2) * Q.ncoarse, Projected->RightFluxes[f][dim]);
    }
  }
  return st;
}

}  // extern "C"

namespace ppm {

// Flux correction of one grid's device state A (cells per slot, dims GD) on
// `stream`; shared by the per-grid and the level entry points.
int correct_refined_fluxes(double *A, long long cells, const int GD[3], hipStream_t stream,
                           const hipPPMParameter *P, const enzo_fluxes *InitialFluxes,
                           const enzo_fluxes *RefinedFluxes,
                           enzo_fluxes *BoundaryFluxesThisTimeStep, enzo_int SUBlingGrid,
                           const enzo_int *GlobalDim, const enzo_int *PeriodicFace,
                           enzo_int CorrectParentBoundaryFlux, const enzo_int *ColourKind) {
  if (!A || !P || !InitialFluxes || !RefinedFluxes) return ENZO_PPM_ERR_ARGUMENT;
  if (CorrectParentBoundaryFlux && !BoundaryFluxesThisTimeStep) return ENZO_PPM_ERR_ARGUMENT;
  if (SUBlingGrid && (!GlobalDim || !PeriodicFace)) return ENZO_PPM_ERR_ARGUMENT;
  if (P->GridRank != 3) return ENZO_PPM_ERR_NOT_SUPPORTED;  // the device holds Vel3 in 3-D only
  const int ncol = (int)P->NumberOfColours;
  for (int c = 0; c < ncol; c++)
    if (!ColourKind || ColourKind[c] < 0 || ColourKind[c] > 2) return ENZO_PPM_ERR_ARGUMENT;
  FluxCorrParams F;
  memset(&F, 0, sizeof(F));
  F.A = A;
  F.cells = cells;
  for (int d = 0; d < 3; d++) F.GD[d] = GD[d];
  F.idual = (int)P->DualEnergyFormalism;
  F.rank = (int)P->GridRank;
  F.tiny_number = P->MinimumPressure;
  // corrected list: density, total energy, gas energy, the velocities in
  // field-number order (Grid_CorrectForRefinedFluxes.C:301-356)
  long long lf[6];
  int nl = 0;
  lf[nl] = P->DensNum; F.list[nl++] = S_D;
  lf[nl] = P->TENum; F.list[nl++] = S_E;
  if (F.idual) { lf[nl] = P->GENum; F.list[nl++] = S_GE; }
  F.nchecked = nl;
  {
    long long vn[3] = {P->Vel1Num, P->Vel2Num, P->Vel3Num};
    int vs[3] = {S_VX, S_VY, S_VZ};
    for (int a = 0; a < 3; a++)
      for (int b = a + 1; b < 3; b++)
        if (vn[b] < vn[a]) { std::swap(vn[a], vn[b]); std::swap(vs[a], vs[b]); }
    for (int a = 0; a < 3; a++) { lf[nl] = vn[a]; F.list[nl++] = vs[a]; }
  }
  F.nl = nl;
  F.cons[F.ncons++] = S_E;
  if (F.idual) F.cons[F.ncons++] = S_GE;
  F.cons[F.ncons++] = S_VX;
  F.cons[F.ncons++] = S_VY;
  F.cons[F.ncons++] = S_VZ;
  for (int c = 0; c < ncol; c++) {
    if (ColourKind[c] == 0) F.spec[F.nspec++] = S_COL0 + c;
    else if (ColourKind[c] == 2) F.cons[F.ncons++] = S_COL0 + c;
  }
  const enzo_fluxes *In = InitialFluxes, *Rf = RefinedFluxes;
  std::vector<double> stage;
  double *dflux = nullptr;
  size_t cap = 0;
  int st = ENZO_PPM_SUCCESS;
  for (int dim = 0; dim < 3 && st == ENZO_PPM_SUCCESS; dim++) {
    if (GD[dim] <= 1) continue;
    // Refined's index ranges of this direction, edited for a SUBling as the
    // reference edits them in place (:147-222)
    long long RLS[3], RLE[3], RRS[3];
    for (int i = 0; i < 3; i++) {
      RLS[i] = Rf->LeftFluxStartGlobalIndex[dim][i];
      RLE[i] = Rf->LeftFluxEndGlobalIndex[dim][i];
      RRS[i] = Rf->RightFluxStartGlobalIndex[dim][i];
    }
    long long RefinedDim[3] = {0, 0, 0}, RefinedOffset[3] = {0, 0, 0}, InitialOffset[3] = {0, 0, 0};
    int corrL = 1, corrR = 1;
    if (!SUBlingGrid) {
      for (int j = 0; j < 3; j++)
        if (In->LeftFluxStartGlobalIndex[dim][j] != RLS[j] || In->LeftFluxEndGlobalIndex[dim][j] != RLE[j])
          st = ENZO_PPM_ERR_ARGUMENT;
    } else {
      const long long G = GlobalDim[dim];
      for (int i = 0; i < 3; i++) RefinedDim[i] = RLE[i] - RLS[i] + 1;
      if (RLS[dim] == 0 && PeriodicFace[2 * dim]) {
        RLS[dim] = G - 1;
        RLE[dim] = G - 1;
      } else {
        RLS[dim]--;
        RLE[dim]--;
      }
      if (RRS[dim] == G - 1 && PeriodicFace[2 * dim + 1]) RRS[dim] = 0;
      else RRS[dim]++;
      corrR = In->RightFluxStartGlobalIndex[dim][dim] == RLS[dim];
      corrL = In->LeftFluxStartGlobalIndex[dim][dim] == RRS[dim];
      for (int i = 0; i < 3; i++)
        RefinedOffset[i] = std::max(In->LeftFluxStartGlobalIndex[dim][i] - RLS[i], 0ll);
      RLS[dim] = 0;
      RefinedOffset[dim] = 0;
    }
    if (st != ENZO_PPM_SUCCESS) break;
    if (!corrL && !corrR) continue;  // a SUBling that meets neither face
    long long Start[3], End[3];
    for (int i = 0; i < 3 && st == ENZO_PPM_SUCCESS; i++) {
      Start[i] = std::max((long long)In->LeftFluxStartGlobalIndex[dim][i], RLS[i]) - P->GridGlobalStart[i];
      End[i] = std::min((long long)In->LeftFluxEndGlobalIndex[dim][i], RLE[i]) - P->GridGlobalStart[i];
      if (Start[i] < 0 || End[i] > GD[i]) st = ENZO_PPM_ERR_ARGUMENT;
    }
    if (st != ENZO_PPM_SUCCESS) break;
    Start[dim] = std::max(Start[dim] - 1, 0ll);
    End[dim] = Start[dim];
    long long Dim[3];
    for (int i = 0; i < 3; i++) Dim[i] = End[i] - Start[i] + 1;
    long long Offset = In->RightFluxStartGlobalIndex[dim][dim] - In->LeftFluxStartGlobalIndex[dim][dim] + 2;
    Offset = std::min(Offset, (long long)GD[dim] - 1);
    if (SUBlingGrid) {
      // this grid's own first / last active layer; Initial spans its whole
      // face (:287-303)
      Offset -= 2;
      Start[dim]++;
      End[dim]++;
      for (int i = 0; i < 3; i++) {
        if (i != dim) {
          Dim[i] = P->GridEndIndex[i] - P->GridStartIndex[i] + 1;
          InitialOffset[i] = std::max(RLS[i] - In->LeftFluxStartGlobalIndex[dim][i], 0ll);
        } else {
          Dim[i] = 1;
          InitialOffset[i] = 0;
        }
      }
    }
    int bndL = 0, bndR = 0;
    if (Start[dim] == P->GridStartIndex[dim] - 1) { bndL = 1; corrL = 0; }
    if (Start[dim] + Offset == P->GridEndIndex[dim] + 1) { bndR = 1; corrR = 0; }
    // the loop region, and the face cells within the grid (both sides)
    long long Ext[3];
    for (int i = 0; i < 3; i++) Ext[i] = End[i] - Start[i] + 1;
    if (Ext[0] <= 0 || Ext[1] <= 0 || Ext[2] <= 0) continue;  // faces do not overlap
    for (int i = 0; i < 3; i++)
      if (Start[i] < 0 || End[i] >= GD[i]) st = ENZO_PPM_ERR_ARGUMENT;
    if (Start[dim] + Offset >= GD[dim] || Start[dim] + Offset < 0) st = ENZO_PPM_ERR_ARGUMENT;
    // every flux index within the arrays: Initial of extent Dim (the
    // grid's own face for a SUBling), Refined of extent RefinedDim
    for (int i = 0; i < 3; i++) {
      const long long inext = In->LeftFluxEndGlobalIndex[dim][i] - In->LeftFluxStartGlobalIndex[dim][i] + 1;
      if (Dim[i] != inext || Ext[i] - 1 + InitialOffset[i] > Dim[i] - 1) st = ENZO_PPM_ERR_ARGUMENT;
      if (SUBlingGrid && Ext[i] - 1 + RefinedOffset[i] > RefinedDim[i] - 1) st = ENZO_PPM_ERR_ARGUMENT;
    }
    if (st != ENZO_PPM_SUCCESS) break;
    const long long nface = Ext[0] * Ext[1] * Ext[2];
    // flux index of loop position (i, j, k) in Initial (and, for a proper
    // subgrid, Refined) and, for a SUBling, in Refined (:455-470)
    auto flux_index = [&](long long i, long long j, long long k) {
      return ((k - Start[2] + InitialOffset[2]) * Dim[1] + (j - Start[1] + InitialOffset[1])) * Dim[0] +
             (i - Start[0] + InitialOffset[0]);
    };
    auto refined_index = [&](long long i, long long j, long long k) {
      if (!SUBlingGrid) return flux_index(i, j, k);
      return ((k - Start[2] + RefinedOffset[2]) * RefinedDim[1] + (j - Start[1] + RefinedOffset[1])) *
                 RefinedDim[0] + (i - Start[0] + RefinedOffset[0]);
    };
    // the grid's own boundary fluxes take the refined ones (the reference
    // assigns them before any undo, which only happens on corrected sides)
    if (CorrectParentBoundaryFlux && (bndL || bndR)) {
      long long GFS[3], GFD[3];
      for (int i = 0; i < 3; i++) {
        GFS[i] = P->GridStartIndex[i];
        GFD[i] = P->GridEndIndex[i] - P->GridStartIndex[i] + 1;
      }
      GFS[dim] = Start[dim];
      GFD[dim] = 1;
      for (int l = 0; l < nl; l++)
        for (long long k = Start[2]; k <= End[2]; k++)
          for (long long j = Start[1]; j <= End[1]; j++)
            for (long long i = Start[0]; i <= End[0]; i++) {
              const long long fi = flux_index(i, j, k);
              const long long gfi = (i - GFS[0]) + (j - GFS[1]) * GFD[0] + (k - GFS[2]) * GFD[1] * GFD[0];
              if (bndL) BoundaryFluxesThisTimeStep->LeftFluxes[lf[l]][dim][gfi] = Rf->LeftFluxes[lf[l]][dim][fi];
              if (bndR) BoundaryFluxesThisTimeStep->RightFluxes[lf[l]][dim][gfi] = Rf->RightFluxes[lf[l]][dim][fi];
            }
    }
    for (int i = 0; i < dim; i++) Offset *= GD[i];
    // stage [list field][side][initial, refined][loop position]: the left
    // side pairs initial left with refined left (proper subgrid) or refined
    // right (SUBling), the right side the other way round
    stage.assign((size_t)nl * 4 * nface, 0.0);
    for (int l = 0; l < nl && st == ENZO_PPM_SUCCESS; l++) {
      const enzo_float *iL = In->LeftFluxes[lf[l]][dim], *iR = In->RightFluxes[lf[l]][dim];
      const enzo_float *rL = Rf->LeftFluxes[lf[l]][dim], *rR = Rf->RightFluxes[lf[l]][dim];
      if (!iL || !iR || !rL || !rR) { st = ENZO_PPM_ERR_ARGUMENT; break; }
      const enzo_float *src[4] = {iL, SUBlingGrid ? rR : rL, iR, SUBlingGrid ? rL : rR};
      double *out = stage.data() + (size_t)l * 4 * nface;
      long long p = 0;
      for (long long k = Start[2]; k <= End[2]; k++)
        for (long long j = Start[1]; j <= End[1]; j++)
          for (long long i = Start[0]; i <= End[0]; i++, p++) {
            const long long fi = flux_index(i, j, k), rfi = refined_index(i, j, k);
            out[p] = src[0][fi];
            out[nface + p] = src[1][rfi];
            out[2 * nface + p] = src[2][fi];
            out[3 * nface + p] = src[3][rfi];
          }
    }
    if (st != ENZO_PPM_SUCCESS) break;
    if (stage.size() > cap) {
      if (dflux) (void)hipFree(dflux);
      dflux = nullptr;
      cap = stage.size();
      if (alloc_dev((void **)&dflux, sizeof(double) * cap)) { st = ENZO_PPM_ERR_HIP; break; }
    }
    if (hipMemcpyAsync(dflux, stage.data(), sizeof(double) * stage.size(), hipMemcpyHostToDevice,
                       stream) != hipSuccess) { st = ENZO_PPM_ERR_HIP; break; }
    for (int i = 0; i < 3; i++) {
      F.Start[i] = (int)Start[i];
      F.Dim[i] = (int)Ext[i];
    }
    F.Offset = Offset;
    F.corr[0] = corrL;
    F.corr[1] = corrR;
    F.subling = SUBlingGrid ? 1 : 0;
    F.flux = dflux;
    F.nface = nface;
    hipLaunchKernelGGL(fluxcorr_kernel, dim3((unsigned)((nface + 255) / 256)), dim3(256), 0,
                       stream, F);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess)
      st = ENZO_PPM_ERR_HIP;
  }
  if (dflux) (void)hipFree(dflux);
  return st;
}

}  // namespace ppm

extern "C" {

int hipPPMCorrectForRefinedFluxes(hipPPMData *D, const hipPPMParameter *P,
                                  const enzo_fluxes *InitialFluxes,
                                  const enzo_fluxes *RefinedFluxes,
                                  enzo_fluxes *BoundaryFluxesThisTimeStep, enzo_int SUBlingGrid,
                                  const enzo_int *GlobalDim, const enzo_int *PeriodicFace,
                                  enzo_int CorrectParentBoundaryFlux, const enzo_int *ColourKind) {
  if (!D) return ENZO_PPM_ERR_ARGUMENT;
  return correct_refined_fluxes(D->A, D->cells, D->GD, D->stream, P, InitialFluxes, RefinedFluxes,
                                BoundaryFluxesThisTimeStep, SUBlingGrid, GlobalDim, PeriodicFace,
                                CorrectParentBoundaryFlux, ColourKind);
}

int hipPPMPeriodicBoundary(hipPPMData *D, const hipPPMParameter *P) {
  if (!D || !P) return ENZO_PPM_ERR_ARGUMENT;
  for (int a = 0; a < P->GridRank; a++) {
    BCParams B;
    B.axis = a;
    B.ghost = (int)P->GridStartIndex[a];
    B.nslot = D->nslot;
    for (int d = 0; d < 3; d++) {
      B.GD[d] = D->GD[d];
      B.act[d] = (int)(P->GridEndIndex[d] - P->GridStartIndex[d] + 1);
    }
    if (!periodic_fits(B, P->GridDimension[a] - P->GridEndIndex[a] - 1)) return ENZO_PPM_ERR_ARGUMENT;
    B.cells = D->cells;
    B.S = D->A;
    long long n = (long long)2 * B.ghost * B.GD[(a + 1) % 3] * B.GD[(a + 2) % 3] * B.nslot;
    hipLaunchKernelGGL(periodic_axis_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       D->stream, B);
    HIPCHK(hipGetLastError());
  }
  return ENZO_PPM_SUCCESS;
}

int hipPPMExternalBoundary(hipPPMData *D, const hipPPMParameter *P, const enzo_int *BoundaryType,
                           const enzo_int *GridOffset, const enzo_int *BoundaryDimension,
                           enzo_int NumberOfBaryonFields, const enzo_float *InflowValue) {
  if (!D || !P || !BoundaryType || !GridOffset || !BoundaryDimension) return ENZO_PPM_ERR_ARGUMENT;
  std::vector<long long> f;
  slot_fields(P, f);
  // check every face first: nothing is written when a face is refused
  for (int pass = 0; pass < 2; pass++)
    for (int a = 0; a < P->GridRank; a++) {
      if (BoundaryDimension[a] <= 1) continue;
      const int S = (int)P->GridStartIndex[a], E = (int)P->GridEndIndex[a], n = E - S + 1;
      for (int face = 0; face < 2; face++) {
        const bool touches = face == 0 ? GridOffset[a] == 0
                                       : GridOffset[a] + D->GD[a] == BoundaryDimension[a];
        if (!touches) continue;
        const int type = (int)BoundaryType[2 * a + face];
        if (type < BK_REFLECTING || type > BK_SHEARING) return ENZO_PPM_ERR_NOT_SUPPORTED;
        const int layers = face == 0 ? S : D->GD[a] - E - 1;
        if (type == BK_PERIODIC || layers <= 0) continue;  // periodic: CopyZonesFromGrid's part
        if (pass == 0) {
          // every source cell inside the grid; shearing reads no ghost of
          // its own face (the reference's in-order loop would see it)
          if (type == BK_REFLECTING && (face == 0 ? 2 * S - 1 >= D->GD[a] : E - (layers - 1) < 0))
            return ENZO_PPM_ERR_ARGUMENT;
          if (type == BK_SHEARING && n < layers) return ENZO_PPM_ERR_NOT_SUPPORTED;
          if (type == BK_INFLOW) {
            if (!InflowValue) return ENZO_PPM_ERR_ARGUMENT;
            for (int s = 0; s < D->nslot; s++)
              if (f[s] >= NumberOfBaryonFields) return ENZO_PPM_ERR_ARGUMENT;
          }
          continue;
        }
        ExtBCParams B;
        memset(&B, 0, sizeof(B));
        B.axis = a;
        B.face = face;
        B.type = type;
        B.layers = layers;
        B.S = S;
        B.E = E;
        B.n = n;
        for (int d = 0; d < 3; d++) B.GD[d] = D->GD[d];
        B.nslot = D->nslot;
        B.vslot = S_VX + a;
        B.cells = D->cells;
        B.A = D->A;
        if (type == BK_INFLOW)
          for (int s = 0; s < D->nslot; s++)
            B.inflow[s] = f[s] >= 0 ? InflowValue[f[s] * 6 + 2 * a + face] : 0.0;
        const long long items = (long long)layers * D->GD[(a + 1) % 3] * D->GD[(a + 2) % 3] * D->nslot;
        // faces in the reference's order (x left, x right, y left, ...):
        // one launch each on the stream
        hipLaunchKernelGGL(ext_bc_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0,
                           D->stream, B);
        HIPCHK(hipGetLastError());
      }
    }
  return ENZO_PPM_SUCCESS;
}

int hipPPMProfile(hipPPMData *D, enzo_int enable) {
  if (!D) return ENZO_PPM_ERR_ARGUMENT;
  D->profile = (int)enable;
  D->ev_used = 0;
  return ENZO_PPM_SUCCESS;
}

int hipPPMProfileRead(hipPPMData *D, enzo_float *total_ms, enzo_long_int *launches) {
  if (!D || !total_ms || !launches) return ENZO_PPM_ERR_ARGUMENT;
  HIPCHK(hipStreamSynchronize(D->stream));
  double tot = 0.0;
  for (size_t k = 0; k + 1 < D->ev_used; k += 2) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, D->ev[k], D->ev[k + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = (enzo_long_int)(D->ev_used / 2);
  D->ev_used = 0;
  return ENZO_PPM_SUCCESS;
}

int hipPPMSynchronize(hipPPMData *D) {
  if (!D) return ENZO_PPM_ERR_ARGUMENT;
  HIPCHK(hipStreamSynchronize(D->stream));
  return ENZO_PPM_SUCCESS;
}

int hipPPMSetLearnedGuess(hipPPMData *D, const float *weights) {
  if (!D) return ENZO_PPM_ERR_ARGUMENT;
  if (!weights) {
    D->learned = false;
    return ENZO_PPM_SUCCESS;
  }
  if (!learn_weights_valid(weights)) return ENZO_PPM_ERR_ARGUMENT;
  if (learn_gate_closed(weights)) {  // no interface can pass: the plain kernels
    D->learned = false;
    return ENZO_PPM_SUCCESS;
  }
  if (!D->learn && alloc_dev((void **)&D->learn, sizeof(float) * kMlpDevFloats))
    return ENZO_PPM_ERR_HIP;
  float dw[kMlpDevFloats];
  learn_device_weights(weights, dw);
  // ordered after the stream's earlier work, visible to its later launches
  HIPCHK(hipStreamSynchronize(D->stream));
  HIPCHK(hipMemcpy(D->learn, dw, sizeof(float) * kMlpDevFloats, hipMemcpyHostToDevice));
  D->learned = true;
  return ENZO_PPM_SUCCESS;
}

int hipPPMSetHarvest(hipPPMData *D, enzo_long_int capacity, enzo_int keep) {
  if (!D || capacity < 0 || keep < 1) return ENZO_PPM_ERR_ARGUMENT;
  HIPCHK(hipStreamSynchronize(D->stream));
  if (D->harvest) (void)hipFree(D->harvest);
  D->harvest = nullptr;
  D->harvest_cap = 0;
  if (capacity == 0) return ENZO_PPM_SUCCESS;
  if (!D->harvest_n && alloc_dev((void **)&D->harvest_n, sizeof(unsigned long long)))
    return ENZO_PPM_ERR_HIP;
  if (alloc_dev((void **)&D->harvest, sizeof(double) * kHarvestDoubles * (size_t)capacity))
    return ENZO_PPM_ERR_HIP;
  HIPCHK(hipMemset(D->harvest_n, 0, sizeof(unsigned long long)));
  D->harvest_cap = capacity;
  D->harvest_keep = (unsigned)keep;
  return ENZO_PPM_SUCCESS;
}

int hipPPMGetHarvest(hipPPMData *D, enzo_float *records, enzo_long_int max_records,
                     enzo_long_int *seen, enzo_long_int *stored) {
  if (!D || max_records < 0 || (max_records > 0 && !records) || !seen || !stored)
    return ENZO_PPM_ERR_ARGUMENT;
  if (D->harvest_cap == 0) {
    *seen = *stored = 0;
    return ENZO_PPM_SUCCESS;
  }
  HIPCHK(hipStreamSynchronize(D->stream));
  unsigned long long n = 0;
  HIPCHK(hipMemcpy(&n, D->harvest_n, sizeof(n), hipMemcpyDeviceToHost));
  const long long kept = std::min<long long>((long long)n, D->harvest_cap);
  const long long m = std::min<long long>(kept, max_records);
  if (m > 0)
    HIPCHK(hipMemcpy(records, D->harvest, sizeof(double) * kHarvestDoubles * (size_t)m,
                     hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(D->harvest_n, 0, sizeof(unsigned long long)));
  *seen = (enzo_long_int)n;
  *stored = (enzo_long_int)m;
  return ENZO_PPM_SUCCESS;
}

#define PPM_STR2(x) #x
#define PPM_STR(x) PPM_STR2(x)
const char *hipPPMVersion(void) {
  return "enzo_ppm 0.2 gfx950 fp64 integers-" PPM_STR(ENZO_PPM_INTEGER_BITS);
}

const char *hipPPMLastError(void) { return g_last_error; }

int hipPPMDeviceCount(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
