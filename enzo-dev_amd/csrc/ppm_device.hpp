// ppm_device.hpp -- device-side pieces shared by the per-grid library
// (enzo_ppm.hip) and the batched AMR-level update (ppm_level.hip): the
// workgroup executor of the sweep tile, the fallback flagging and the
// per-group fallback body, the fused pressure + Courant cell, and the
// per-sweep preparation item.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "ppm_sweep.hpp"

namespace ppm {

// A workgroup-uniform int as such (the compiler then keeps what derives
// from it in scalar registers), and a per-work-item int the compiler must
// treat as new at this point (no code that depends on it is hoisted above
// it).  The CPU execution of this source (tests/emu/hipcpu) runs work-items
// one by one: there both are the value itself.
#ifdef HIPCPU_RUNTIME
#define PPM_UNIFORM(x) (x)
#define PPM_OPAQUE_VGPR(x) ((void)(x))
#else
#define PPM_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#define PPM_OPAQUE_VGPR(x) asm volatile("" : "+v"(x))
#endif

// LDS of the executor's votes, one per kernel (a local __shared__ array in
// the member templates below would be one per instantiation)
__device__ inline int *exec_lds() {
  __shared__ int slots[17];  // [0, 16): a count per wavefront, [16]: first()
  return slots;
}

// Workgroup executor of the tile phases (ppm_sweep.hpp): run = phase then
// barrier; count = number of work-items with a nonzero result; first = the
// result of work-item 0, broadcast.
struct DevExec {
  Regs r;
  int tid = (int)threadIdx.x;  // the work-item's index passed to the phases
  template <class F>
  __device__ void run(F f) {
    f(tid, r);
    __syncthreads();
  }
  // a wavefront vote per wavefront, its population count in LDS, the sum
  // over the workgroup's wavefronts (at most 16: 1024 work-items)
  template <class F>
  __device__ int count(F f) {
    int *part = exec_lds();
    int v = f(tid, r);
    const unsigned long long b = __ballot(v != 0);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    int n = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) n += part[w];
    __syncthreads();  // part is free for the next call
    return n;
  }
  template <class F>
  __device__ int first(F f) {
    int &v0 = exec_lds()[16];
    int v = f(tid, r);
    if (threadIdx.x == 0) v0 = v;
    __syncthreads();
    v = v0;
    __syncthreads();  // v0 is free for the next call
    return v;
  }
};

// One-pencil fallback workgroups: the whole pencil in LDS, one work-item per
// interface, in workgroups of 128, 256, 512 or kFbTS = 1024 work-items (the
// narrowest that holds every pencil of the launch: fewer registers spilled
// and more workgroups per CU than the widest).
constexpr int kFbTS = 1024;
constexpr int kFbWidths[4] = {128, 256, 512, 1024};
static_assert(kFbWidths[3] == kFbTS, "the widest one-pencil workgroup is kFbTS");
// The per-cell fallback launches kFbBlocks * kFbTS work-items in all (with
// no flagged group each workgroup only reads the count); RiemannSolver HLL /
// HLLC send every group there, kFbBlocksWhole * kFbTS work-items: a CU's
// worth of LDS per 1024 work-items.
constexpr int kFbBlocks = 64;
constexpr int kFbBlocksWhole = 256;

// narrowest one-pencil width whose tile holds the sweep axis of S (0: none)
PPM_HD bool pencil_fits_width(const SweepParams &S, int ts) {
  return ts == 128 ? pencil_fits<128>(S) : ts == 256 ? pencil_fits<256>(S)
       : ts == 512 ? pencil_fits<512>(S) : ts == 1024 ? pencil_fits<1024>(S) : false;
}
inline int fb_width_of(const SweepParams &S) {
  for (int ts : kFbWidths)
    if (pencil_fits_width(S, ts)) return ts;
  return 0;
}

// f(std::integral_constant<int, TS>()) for the one-pencil width ts, which
// must be one of kFbWidths (each is instantiated here, so a width that is
// not in the list cannot reach a narrower kernel)
template <class F>
inline hipError_t with_fb_width(int ts, F f) {
  switch (ts) {
    case 128: f(std::integral_constant<int, 128>()); break;
    case 256: f(std::integral_constant<int, 256>()); break;
    case 512: f(std::integral_constant<int, 512>()); break;
    case 1024: f(std::integral_constant<int, 1024>()); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
// workgroups of width ts for a launch of `total` work-items
inline unsigned fb_blocks(int ts, int total) { return (unsigned)(total / ts); }
using FbCfg = Cfg<1, kFbTS>;

// Work-item 0 of a sweep workgroup whose tile saw a failing cell or an
// error: put the tile's pencil group on the fallback list once (the group
// flag deduplicates the tiles of one group); the entry carries P.fb_tag
// (the grid of a batched level, 0 otherwise).
__device__ inline void flag_group(const SweepParams &P, int grp) {
  if (atomicExch(&P.fb_flag[grp], 1) == 0)
    P.fb_list[atomicAdd(P.fb_count, 1u)] = P.fb_tag + (unsigned)grp;
}

// Every pencil of group `grp` (P.tp pencils along t1) redone whole with the
// per-cell fallback, then the group flag cleared.
template <int UNI, int LEARN = 0, int TS = kFbTS>
__device__ inline void fallback_group(const SweepParams &P, int grp, double *sm, DevExec &ex) {
  const int tb = grp % P.nblk_t1, t2 = P.t2_0 + grp / P.nblk_t1;
  for (int p = 0; p < P.tp && tb * P.tp + p < P.n_t1; p++) {
    if (!P.fb_ok) {
      if (threadIdx.x == 0) atomicAdd(&P.counters[3], 1ull);
      break;
    }
    TileCounts tc = sweep_pencil_fb<TS, UNI, DevExec, LEARN>(P, tb * P.tp + p, t2, sm, ex);
    if (threadIdx.x == 0) {
      if (tc.nfb) atomicAdd(&P.counters[0], (unsigned long long)tc.nfb);
      if (tc.err) atomicAdd(&P.counters[2], (unsigned long long)tc.err);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) P.fb_flag[grp] = 0;
}

// End of a fallback launch: the last workgroup to finish (every workgroup has
// read the list length by then) clears the length for the next launch.
__device__ inline void fallback_launch_done(unsigned *count, unsigned *done) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      *count = 0u;
      *done = 0u;
      __threadfence();
    }
  }
}

// ------------------------------------------------------------------------
// Fused pressure + Courant condition (Grid_ComputePressure.C:96-166,
// calc_dt.F:129-177).  Active cells contribute
// dt1 = aye / sum_d (c_s + |v_d|)/dx_d to a device-wide minimum kept as the
// bit pattern of a positive double (ordered like the value).
struct PresParams {
  int rank, idual;
  int GD[3], st[3], en[3];
  long long cells;
  const double *A;
  double *pres;
  double gamma, tiny_number, min_support;
  const double *cw[3];  // CellWidth per axis (calc_dt widths)
  double aye;
  int want_dt;
  unsigned long long *dtmin;
};

__device__ inline double pressure_of(const PresParams &Q, long long g, double min_support) {
  const auto A = as_global(Q.A);
  double density = A[S_D * Q.cells + g];
  double gas_energy;
  if (Q.idual == 0) {
    double v1 = A[S_VX * Q.cells + g];
    double v2 = Q.rank > 1 ? A[S_VY * Q.cells + g] : 0.0;
    double v3 = Q.rank > 2 ? A[S_VZ * Q.cells + g] : 0.0;
    gas_energy = A[S_E * Q.cells + g] - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
  } else {
    gas_energy = A[S_GE * Q.cells + g];
  }
  double p = (Q.gamma - 1.0) * density * gas_energy;
  if (p < Q.tiny_number) p = Q.tiny_number;
  double min_pressure = min_support * (Q.gamma - 1.0) * density * density;
  if (p < min_pressure) p = min_pressure;
  return p;
}

// pressure of cell g into Q.pres; returns the cell's Courant step (1e20 for
// ghost cells or without the time step)
__device__ inline double pressure_dt_cell(const PresParams &Q, long long g) {
  double p = pressure_of(Q, g, Q.min_support);
  as_global(Q.pres)[g] = p;
  if (!Q.want_dt) return 1.0e20;
  int i = (int)(g % Q.GD[0]);
  long long r = g / Q.GD[0];
  int j = (int)(r % Q.GD[1]);
  int k = (int)(r / Q.GD[1]);
  if (i < Q.st[0] || i > Q.en[0] || j < Q.st[1] || j > Q.en[1] || k < Q.st[2] || k > Q.en[2])
    return 1.0e20;
  // calc_dt's pressure uses minimum-support coefficient 0
  double pdt = Q.min_support != 0.0 ? pressure_of(Q, g, 0.0) : p;
  const auto A = as_global(Q.A);
  const auto cw0 = as_global(Q.cw[0]), cw1 = as_global(Q.cw[1]), cw2 = as_global(Q.cw[2]);
  double d = A[S_D * Q.cells + g];
  double cs = fmx(sqrt(Q.gamma * pdt / d), kTiny);
  double u = A[S_VX * Q.cells + g];
  if (Q.rank == 1) return cw0[i] * Q.aye / (cs + fabs(u - 0.0));
  if (Q.rank == 2) {
    double v = A[S_VY * Q.cells + g];
    return Q.aye / ((cs + fabs(u - 0.0)) / cw0[i] + (cs + fabs(v - 0.0)) / cw1[j]);
  }
  double v = A[S_VY * Q.cells + g], w = A[S_VZ * Q.cells + g];
  return Q.aye / ((cs + fabs(u - 0.0)) / cw0[i] + (cs + fabs(v - 0.0)) / cw1[j] +
                  (cs + fabs(w - 0.0)) / cw2[k]);
}

// wavefront then workgroup minimum of `local` into *dtmin (blockDim a
// multiple of 64, at most 1024)
__device__ inline void block_min_dt(double local, unsigned long long *dtmin) {
  for (int off = 32; off > 0; off >>= 1) local = fmn(local, __shfl_xor(local, off, 64));
  __shared__ double wmin[16];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmn(m, wmin[w]);
    // the device-wide minimum only decreases: a workgroup whose minimum is
    // not below the value it reads skips the atomic (one address, updated
    // by every workgroup of the launch, serialises them)
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    if (bits < __hip_atomic_load(dtmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(dtmin, bits);
  }
}

// ------------------------------------------------------------------------
// Periodic ghost fill along one axis (the periodic case of
// ExternalBoundary::SetExternalBoundary): item gid copies one ghost cell of
// one slot from the opposite side of the active region.
struct BCParams {
  int axis, ghost, nslot;
  int GD[3], act[3];
  long long cells;
  double *S;
};

// Whether the periodic fill of B's axis is defined and fits its items:
// as many ghost cells on the right (`right_ghost`) as on the left, an
// active extent of at least the ghost count (each ghost cell is then one
// period from an active cell, the images CopyZonesFromGrid takes; a
// narrower grid would also read ghost cells the same launch writes), and
// fewer than 2^31 items.
inline bool periodic_fits(const BCParams &B, long long right_ghost) {
  const int a = B.axis;
  const long long n = 2LL * B.ghost * B.GD[(a + 1) % 3] * B.GD[(a + 2) % 3] * B.nslot;
  return right_ghost == B.ghost && B.ghost >= 0 && B.act[a] >= B.ghost && n < (1LL << 31);
}

__device__ inline long long periodic_items(const BCParams &B) {
  return (long long)2 * B.ghost * B.GD[(B.axis + 1) % 3] * B.GD[(B.axis + 2) % 3] * B.nslot;
}

// Item gid walks the box of the axis' ghost layers (extent 2 * ghost along
// the axis, the grid's extent across it) with x fastest, so that the loads
// and stores of a wavefront are contiguous along x (runs of `ghost` cells
// when the axis is x).  Layer l < ghost is left ghost cell l, taken from
// active cell l + act; layer l >= ghost is right ghost cell l + act, taken
// from active cell l (act >= ghost, checked on the host).  The host also
// keeps the item count below 2^31, so 32-bit arithmetic.
__device__ inline void periodic_item(const BCParams &B, int gid) {
  const int a = B.axis, g = B.ghost, act = B.act[a];
  const int n0 = a == 0 ? 2 * g : B.GD[0], n1 = a == 1 ? 2 * g : B.GD[1],
            n2 = a == 2 ? 2 * g : B.GD[2];
  const int per_slot = n0 * n1 * n2;
  const int slot = gid / per_slot;
  int r = gid - slot * per_slot;
  const int x0 = r % n0;
  r /= n0;
  const int x1 = r % n1, x2 = r / n1;
  const int l = a == 0 ? x0 : a == 1 ? x1 : x2;
  const int sa = a == 0 ? 1 : a == 1 ? B.GD[0] : B.GD[0] * B.GD[1];  // stride of the axis
  const int base = (x2 * B.GD[1] + x1) * B.GD[0] + x0 - l * sa;     // cell 0 of the line
  const int dst = l < g ? l : l + act, src = l < g ? l + act : l;
  double *S = B.S + (long long)slot * B.cells;
  S[base + dst * sa] = S[base + src * sa];
}

// ------------------------------------------------------------------------
// External boundary of one domain face (ExternalBoundary::SetExternalBoundary,
// ExternalBoundary_SetExternalBoundaryIO.C:50-1010 without OOC_BOUNDARY):
// ghost layer i of the face takes, per its boundary type, the mirrored
// active cell (reflecting; the normal velocity's sign flipped), the edge
// cell (outflow), the face's value (inflow) or the cell one active extent
// inwards (shearing).  One work-item per (slot, ghost cell of the face).
enum BoundaryKind { BK_REFLECTING = 0, BK_OUTFLOW = 1, BK_INFLOW = 2, BK_PERIODIC = 3, BK_SHEARING = 4 };
struct ExtBCParams {
  int axis, face, type, layers;  // face 0 left, 1 right; ghost layers of the face
  int S, E, n;                   // GridStartIndex, GridEndIndex, active cells along axis
  int GD[3];
  int nslot, vslot;              // slots; the normal velocity's slot
  long long cells;
  double *A;
  double inflow[S_COL0 + ENZO_PPM_MAX_COLOR];  // per slot (inflow faces)
};

__device__ inline long long ext_bc_items(const ExtBCParams &B) {
  return (long long)B.layers * B.GD[(B.axis + 1) % 3] * B.GD[(B.axis + 2) % 3] * B.nslot;
}

__device__ inline void ext_bc_item(const ExtBCParams &B, long long gid) {
  const int a = B.axis, b = (a + 1) % 3, c = (a + 2) % 3;
  const long long per_slot = (long long)B.layers * B.GD[b] * B.GD[c];
  const int slot = (int)(gid / per_slot);
  long long r = gid % per_slot;
  const int i = (int)(r % B.layers);
  r /= B.layers;
  int x[3];
  x[b] = (int)(r % B.GD[b]);
  x[c] = (int)(r / B.GD[b]);
  int dst, src;
  if (B.face == 0) {
    dst = i;
    src = B.type == BK_REFLECTING ? 2 * B.S - 1 - i : B.type == BK_OUTFLOW ? B.S : i + B.n;
  } else {
    dst = B.E + 1 + i;
    src = B.type == BK_REFLECTING ? B.E - i : B.type == BK_OUTFLOW ? B.E : dst - B.n;
  }
  double *F = B.A + (long long)slot * B.cells;
  x[a] = dst;
  const long long gd = ((long long)x[2] * B.GD[1] + x[1]) * B.GD[0] + x[0];
  double v;
  if (B.type == BK_INFLOW) {
    v = B.inflow[slot];
  } else {
    x[a] = src;
    v = F[((long long)x[2] * B.GD[1] + x[1]) * B.GD[0] + x[0]];
    if (B.type == BK_REFLECTING) v = (slot == B.vslot ? -1.0 : 1.0) * v;
  }
  F[gd] = v;
}

// ------------------------------------------------------------------------
// Comoving expansion terms of one cell (Grid_ComovingExpansionTerms.C:55-130
// with expand_terms.F, ENERGY_METHOD3 / VELOCITY_METHOD3, HydroMethod 0):
// the pressure at the interpolation weight `coef` between the state kept at
// the step's start (`old`, slots S_D, S_E, S_VX..S_VZ, S_GE as in A) and the
// current state (Grid_ComputePressure.C:56-61, :170-240), or the current
// state's pressure when `interp` is 0 (its time == Time branch); then the
// semi-implicit decay of gas energy, total energy and velocities with
// Coefficient = dt (da/dt) / a.  Every cell, ghosts included.
struct ExpandParams {
  long long cells;
  int rank, idual, interp;
  double gamma, tiny_number, coef, Coefficient;
  double *A;    // current state (slots as in the sweeps)
  double *old;  // state at the step's start, slots S_D..S_GE (interp only)
};

// The slots of the step's starting state that expand_cell reads (the
// interpolated pressure's inputs): density and gas energy with dual energy,
// density, total energy and the rank's velocities without it.  Returns how
// many, in s[] (at most 5).
PPM_HD int expansion_old_slots(int idual, int rank, int s[5]) {
  int n = 0;
  s[n++] = S_D;
  if (idual) {
    s[n++] = S_GE;
  } else {
    s[n++] = S_E;
    for (int d = 0; d < rank; d++) s[n++] = S_VX + d;
  }
  return n;
}

__device__ inline void expand_cell(const ExpandParams &X, long long g) {
  const auto A = as_global(X.A);
  const long long n = X.cells;
  // every field of the cell read before any is written, so that its loads
  // are in flight together (each field is written once, from these values)
  const double d = A[S_D * n + g], e0 = A[S_E * n + g], v1 = A[S_VX * n + g];
  const double v2 = X.rank > 1 ? A[S_VY * n + g] : 0.0;
  const double v3 = X.rank > 2 ? A[S_VZ * n + g] : 0.0;
  const double ge0 = X.idual != 0 ? A[S_GE * n + g] : 0.0;
  double density, gas_energy;
  if (X.interp) {
    const auto O = as_global(X.old);
    const double c = X.coef, co = 1 - X.coef;
    density = c * d + co * O[S_D * n + g];
    if (X.idual == 0) {
      const double total_energy = c * e0 + co * O[S_E * n + g];
      const double velocity1 = c * v1 + co * O[S_VX * n + g];
      double velocity2 = 0, velocity3 = 0;
      if (X.rank > 1) velocity2 = c * v2 + co * O[S_VY * n + g];
      if (X.rank > 2) velocity3 = c * v3 + co * O[S_VZ * n + g];
      const double kineticE =
          0.5 * (velocity1 * velocity1 + velocity2 * velocity2 + velocity3 * velocity3);
      gas_energy = total_energy - kineticE;
    } else {
      gas_energy = c * ge0 + co * O[S_GE * n + g];
    }
  } else {
    density = d;
    gas_energy = X.idual == 0 ? e0 - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3) : ge0;
  }
  double p = (X.gamma - 1.0) * density * gas_energy;
  if (p < X.tiny_number) p = X.tiny_number;
  // expand_terms.F:66-166
  const double cf = X.Coefficient;
  if (X.idual == 1) A[S_GE * n + g] = ge0 * (1.0 - cf) / (1.0 + cf);
  double e = e0 * (1.0 - cf) / (1.0 + cf);
  e = fmx(e - cf * (3.0 - 2.0 / (X.gamma - 1.0)) * p / d, 0.5 * e);
  A[S_E * n + g] = e;
  A[S_VX * n + g] = v1 * (1.0 - 0.5 * cf) / (1.0 + 0.5 * cf);
  if (X.rank > 1) A[S_VY * n + g] = v2 * (1.0 - 0.5 * cf) / (1.0 + 0.5 * cf);
  if (X.rank > 2) A[S_VZ * n + g] = v3 * (1.0 - 0.5 * cf) / (1.0 + 0.5 * cf);
}

// ------------------------------------------------------------------------
// AMR flux correction of the cells just outside one face pair of a proper
// subgrid, or of this grid's own first / last active layer for a SUBling
// (Grid_CorrectForRefinedFluxes.C:53-870, Cartesian, PPM direct Euler,
// RadiativeCooling 0, FluxCorrection 1).  One
// work-item per face position; it treats the left cell and then the right
// cell, each in the reference's order: conserved form (energies, velocities
// and kind-2 colours times density, species-like colours over density),
// the corrections in list order (density, total energy, gas energy, the
// velocities) with the undo of all of them when density or an energy would
// not stay positive, back to specific form, dual-energy consistency, and
// species times the new density.
struct FluxCorrParams {
  double *A;
  long long cells;
  int GD[3];
  int Start[3], Dim[3];
  long long Offset;           // right cell = left cell + Offset
  int corr[2];                // correct the left / right cells
  int nl, nchecked;           // corrected fields; the first nchecked must stay > 0
  int list[6];                // their slots
  int ncons, nspec;
  int cons[6 + 45], spec[45]; // slots multiplied by / divided by density
  int idual, rank;
  double tiny_number;
  const double *flux;         // [field in list][side][initial, refined][face position]
  long long nface;
  int subling;                // SUBlingGrid: the pairs are (initial left, refined right)
                              // and (initial right, refined left), signs swapped
};

__device__ inline void fluxcorr_side(const FluxCorrParams &F, long long x, int side, long long fi) {
  double *A = F.A;
  const long long n = F.cells;
  for (int c = 0; c < F.ncons; c++) A[F.cons[c] * n + x] *= A[S_D * n + x];
  for (int c = 0; c < F.nspec; c++) A[F.spec[c] * n + x] /= A[S_D * n + x];
  if (F.corr[side]) {
    double ini[6], ref[6];
    for (int l = 0; l < F.nl; l++) {
      const double *q = F.flux + ((long long)(l * 2 + side) * 2) * F.nface;
      ini[l] = q[fi];
      ref[l] = q[F.nface + fi];
    }
    // a proper subgrid adds initial - refined on the left and its negative
    // on the right; a SUBling the other way round
    // (Grid_CorrectForRefinedFluxes.C:491-544)
    const bool neg = (side == 1) != (F.subling != 0);
    // the reference's undo on a SUBling's right face subtracts the amount of
    // the field found non-positive from every field corrected before it
    // (:638-652)
    const bool own_amount = F.subling != 0 && side == 1;
    for (int l = 0; l < F.nl; l++) {
      double &v = A[F.list[l] * n + x];
      if (neg) v += -(ini[l] - ref[l]);
      else v += ini[l] - ref[l];
      if (l < F.nchecked && v <= 0) {
        const double amt = ini[l] - ref[l];
        for (int u = 0; u < F.nl; u++) {
          if (u <= l) {
            if (own_amount) A[F.list[u] * n + x] -= amt;
            else if (neg) A[F.list[u] * n + x] -= -(ini[u] - ref[u]);
            else A[F.list[u] * n + x] -= (ini[u] - ref[u]);
          }
          ref[u] = ini[u];
        }
      }
    }
  }
  for (int c = 0; c < F.ncons; c++) A[F.cons[c] * n + x] /= A[S_D * n + x];
  if (F.idual) {
    double ge = fmx(A[S_GE * n + x], F.tiny_number);
    A[S_GE * n + x] = ge;
    const double v1 = A[S_VX * n + x];
    double te = ge + 0.5 * v1 * v1;
    if (F.rank > 1) {
      const double v2 = A[S_VY * n + x];
      te += 0.5 * v2 * v2;
    }
    if (F.rank > 2) {
      const double v3 = A[S_VZ * n + x];
      te += 0.5 * v3 * v3;
    }
    A[S_E * n + x] = te;
  }
  for (int c = 0; c < F.nspec; c++) A[F.spec[c] * n + x] *= A[S_D * n + x];
}

// ------------------------------------------------------------------------
// Per-sweep preparation item `gid`: qc(i) = dt/dx(i) for gid < N (used by
// the fluxes, flux_twoshock.F:366, and the gas-energy source, :393), then the
// flattening of line 1 of every slice, one (slice, cell in [i1-1, i2+1]) per
// item.  The reference hands `flatten` (line 1) to intvar for every line of
// the slice (inteuler.F:241-289), so one coefficient row per slice is all the
// sweep needs.  Items: N + (i2 - i1 + 3) * n_outer.
__device__ inline void prep_item(const FlatParams &F, long long gid) {
  if (gid < F.N) {
    const double dt = F.dtp ? *as_global(F.dtp) : F.dt;
    as_global(F.qc)[gid] = dt / as_global(F.dx)[gid];
    return;
  }
  if (!F.iflatten) return;
  const long long k = gid - F.N;
  const int ncell = F.i2 - F.i1 + 3;
  if (k >= (long long)ncell * F.n_outer) return;
  const int outer = (int)(k / ncell);
  const int cell = F.i1 - 1 + (int)(k % ncell);
  as_global(F.flat)[(long long)outer * F.N + (cell - 1)] = flat_value(F, outer, cell);
}


// ------------------------------------------------------------------------
// Host: flux correction of one grid's device state (A: the grid's slots,
// `cells` apart; GD: its dimensions) on `stream`, as
// hipPPMCorrectForRefinedFluxes describes; waits for the stream.
int correct_refined_fluxes(double *A, long long cells, const int GD[3], hipStream_t stream,
                           const hipPPMParameter *P, const enzo_fluxes *InitialFluxes,
                           const enzo_fluxes *RefinedFluxes,
                           enzo_fluxes *BoundaryFluxesThisTimeStep, enzo_int SUBlingGrid,
                           const enzo_int *GlobalDim, const enzo_int *PeriodicFace,
                           enzo_int CorrectParentBoundaryFlux, const enzo_int *ColourKind);

// ------------------------------------------------------------------------
// Host: BaryonField number of state slot s, or -1 (a slot the grid does
// not hold)
inline long long slot_field(const hipPPMParameter *P, int s) {
  if (s == S_D) return P->DensNum;
  if (s == S_E) return P->TENum;
  if (s == S_VX) return P->Vel1Num;
  if (s == S_VY) return P->GridRank > 1 ? P->Vel2Num : -1;
  if (s == S_VZ) return P->GridRank > 2 ? P->Vel3Num : -1;
  if (s == S_GE) return P->DualEnergyFormalism ? P->GENum : -1;
  return s - S_COL0 < P->NumberOfColours ? P->colnum[s - S_COL0] : -1;
}

// Host: Enzo field number of each device slot (-1: none)
inline void slot_fields(const hipPPMParameter *P, std::vector<long long> &f) {
  f.assign(S_COL0 + (int)P->NumberOfColours, -1);
  f[S_D] = P->DensNum;
  f[S_E] = P->TENum;
  f[S_VX] = P->Vel1Num;
  if (P->GridRank > 1) f[S_VY] = P->Vel2Num;
  if (P->GridRank > 2) f[S_VZ] = P->Vel3Num;
  if (P->DualEnergyFormalism) f[S_GE] = P->GENum;
  for (int n = 0; n < P->NumberOfColours; n++) f[S_COL0 + n] = P->colnum[n];
}

// ------------------------------------------------------------------------
// Host: the face records of one sweep axis for a grid's SubgridFluxes
// (Grid_[xyz]EulerSweep.C:313-376 capture positions): one record per
// subgrid face crossing the axis, sorted by interface position, with
// staging offsets of nfld fields each, and head[pos] = first record at or
// after interface pos.  ENZO_PPM_ERR_ARGUMENT for a face outside the grid.
inline int build_face_records(const int GD[3], const hipPPMParameter *P,
                              enzo_fluxes *const *SubgridFluxes, int nsub,
                              const enzo_long_int *GridGlobalStart, int nfld, int dim,
                              std::vector<FaceRec> &sorted, std::vector<int> &sub,
                              std::vector<int> &side, std::vector<int> &head, long long &total) {
  sorted.clear();
  sub.clear();
  side.clear();
  head.clear();
  total = 0;
  if (dim >= P->GridRank) return ENZO_PPM_SUCCESS;
  const int N = GD[dim];
  const int i1 = (int)P->GridStartIndex[dim] + 1, i2 = (int)P->GridEndIndex[dim] + 1;
  const int ai = dim == 0 ? 1 : 0, aj = dim == 2 ? 1 : 2;
  std::vector<FaceRec> recs;
  std::vector<int> rsub, rside;
  std::vector<std::pair<int, int>> order;  // (pos, record)
  for (int n = 0; n < nsub; n++) {
    const enzo_fluxes *F = SubgridFluxes[n];
    if (!F) return ENZO_PPM_ERR_ARGUMENT;
    for (int sd = 0; sd < 2; sd++) {
      FaceRec R;
      memset(&R, 0, sizeof(R));
      long long face = sd == 0 ? F->LeftFluxStartGlobalIndex[dim][dim] - GridGlobalStart[dim]
                               : F->RightFluxStartGlobalIndex[dim][dim] - GridGlobalStart[dim] + 1;
      R.pos = (int)face + 1;  // 1-based interface at the left edge of cell `face`
      R.lo_i = (int)(F->RightFluxStartGlobalIndex[dim][ai] - GridGlobalStart[ai]);
      R.hi_i = (int)(F->RightFluxEndGlobalIndex[dim][ai] - GridGlobalStart[ai]);
      R.lo_j = (int)(F->RightFluxStartGlobalIndex[dim][aj] - GridGlobalStart[aj]);
      R.hi_j = (int)(F->RightFluxEndGlobalIndex[dim][aj] - GridGlobalStart[aj]);
      if (R.pos < i1 || R.pos > i2 + 1 || R.hi_i < R.lo_i || R.hi_j < R.lo_j)
        return ENZO_PPM_ERR_ARGUMENT;
      R.nface = (long long)(R.hi_i - R.lo_i + 1) * (R.hi_j - R.lo_j + 1);
      order.push_back({R.pos, (int)recs.size()});
      recs.push_back(R);
      rsub.push_back(n);
      rside.push_back(sd);
    }
  }
  std::stable_sort(order.begin(), order.end());
  for (auto &pr : order) {
    FaceRec R = recs[pr.second];
    R.base = total;
    total += R.nface * nfld;
    sorted.push_back(R);
    sub.push_back(rsub[pr.second]);
    side.push_back(rside[pr.second]);
  }
  head.assign(N + 2, 0);
  size_t r = 0;
  for (int pos = 0; pos <= N + 1; pos++) {
    while (r < sorted.size() && sorted[r].pos < pos) r++;
    head[pos] = (int)r;
  }
  return ENZO_PPM_SUCCESS;
}

// Host: a grid's captured fluxes of one axis (staging `h`) into its
// SubgridFluxes, with the velocity guards of Grid_[xyz]EulerSweep.C:343-357.
inline void scatter_face_fluxes(const std::vector<double> &h, const std::vector<FaceRec> &recs,
                                const std::vector<int> &sub, const std::vector<int> &side,
                                const std::vector<long long> &slot_field, const long long nact[3],
                                int dim, enzo_fluxes *const *SubgridFluxes) {
  for (size_t r = 0; r < recs.size(); r++) {
    const FaceRec &R = recs[r];
    enzo_fluxes *F = SubgridFluxes[sub[r]];
    for (int s = 0; s < (int)slot_field.size(); s++) {
      if (slot_field[s] < 0) continue;
      if (s >= S_VX && s <= S_VZ && (s - S_VX) != dim && nact[s - S_VX] <= 1) continue;
      enzo_float *out = side[r] == 0 ? F->LeftFluxes[slot_field[s]][dim] : F->RightFluxes[slot_field[s]][dim];
      if (!out) continue;
      memcpy(out, h.data() + R.base + (size_t)s * R.nface, sizeof(double) * R.nface);
    }
  }
}

}  // namespace ppm
