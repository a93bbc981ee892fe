// ppm_sweep.hpp -- body of the fused PPM-DE directional sweep for CDNA4.
//
// One workgroup updates a tile of TP pencils x one segment of the sweep axis.
// Each thread owns one interface (Riemann problem) of one pencil; the pencil
// segment and its ghost cells sit in LDS.  Stages, in order:
//   load -> characteristic speeds / density second differences -> steepening
//   -> per variable (d,p,u,v,w,ge) PPM slopes, parabolae and interface
//   averages -> Eulerian corrections, two-shock Newton, Godunov fluxes ->
//   negative-state check -> conservative update -> dual-energy sync -> write
//   back + face-flux capture -> per colour: reconstruction, flux, update.
// The arithmetic of every stage follows the reference Fortran expression by
// expression (no contraction, IEEE division and square root), so that the
// limiter and dual-energy branches take the same sides:
//   calcdiss.F:193-442, inteuler.F:140-551, intvar.F:82-174,
//   twoshock.F:141-254, flux_twoshock.F:84-456, euler.F:121-279,
//   pgas2d_dual.F:70-118.
//
// The body is written as phase functions of the thread id; the caller puts
// a workgroup barrier between phases.  The same code runs as the HIP kernel
// and, for index and arithmetic checks on a CPU, in the block emulator of
// tests/emu (compiled with PPM_EMU_CHECKS, which bounds-checks every LDS and
// global access).
#pragma once

#include <math.h>

#include <cmath>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PPM_HD __host__ __device__ inline
#else
#define PPM_HD inline
#endif

#ifdef PPM_EMU_CHECKS
#include <assert.h>
#define PPM_ASSERT(...) assert((__VA_ARGS__))
#else
#define PPM_ASSERT(...) ((void)0)
#endif

#include "ppm_learn.hpp"

// atomic add of the harvest counter: device atomics in the HIP and CPU-HIP
// builds, a plain add in the block emulator (one work-item at a time)
#if defined(__HIPCC__) || defined(HIPCPU_RUNTIME)
#define PPM_ATOMIC_ADD_U64(p, v) atomicAdd((p), (unsigned long long)(v))
#else
inline unsigned long long ppm_add_u64(unsigned long long *p, unsigned long long v) {
  unsigned long long o = *p;
  *p = o + v;
  return o;
}
#define PPM_ATOMIC_ADD_U64(p, v) ppm_add_u64((p), (unsigned long long)(v))
#endif

namespace ppm {

constexpr double kTiny = 1e-20;        // fortran_types.def:7-12
constexpr double kColorFloor = 1e-35;  // fortran.def:8
constexpr double kFt = 4.0 / 3.0;      // intvar.F:59, inteuler.F:133

PPM_HD double fmx(double a, double b) { return a > b ? a : b; }  // Fortran MAX as flang lowers it
PPM_HD double fmn(double a, double b) { return a < b ? a : b; }  // Fortran MIN
PPM_HD double sgn1(double b) { return copysign(1.0, b); }       // SIGN(1,b)

// field slots of the device state (one array of `cells` doubles each)
enum Slot { S_D = 0, S_E = 1, S_VX = 2, S_VY = 3, S_VZ = 4, S_GE = 5, S_COL0 = 6 };

// per-axis coefficient tables, each N doubles, index = cell-1
enum Tab {
  T_C1 = 0, T_C2, T_C3, T_C4, T_C5, T_C6, T_DX2I, T_DX, T_SDXP, T_SDXM, T_QA3,
  T_STNUM, T_STDEN, T_DIFFX, T_DDEN, NTAB
};

struct FaceRec {          // one subgrid face crossing the sweep axis
  int32_t pos;            // 1-based interface index along the sweep
  int32_t lo_i, hi_i;     // range along the lower transverse axis
  int32_t lo_j, hi_j;     // range along the higher transverse axis
  int32_t pad;
  int64_t base;           // offset of this face's block in the staging buffer
  int64_t nface;          // (hi_i-lo_i+1)*(hi_j-lo_j+1)
};

struct SweepParams {
  int dim, N, i1, i2, rank;
  int GD[3];
  long long cells;
  int ax_t1, ax_t2;       // grid axes enumerated by t1 (within a block) and t2
  int st_s, st_1, st_2;   // grid strides of the sweep, t1 and t2 axes (set_strides)
  int n_t1, n_t2, t2_0;   // pencil counts; first t2 (slab-serial launches)
  int nseg, nblk_t1;
  const int *seg_o, *seg_w;  // owned cells of each segment: [o, o+w)
  const double *A;        // source state  (slot*cells + index)
  double *B;              // destination state
  int ncol;
  const double *pres;     // stale pressure
  const double *gr;       // acceleration along dim (gravity) or null
  const double *flat;     // flattening of line 1 of each slice: [outer][N]
  const double *tab, *tab_row, *tab_out;  // tables of sweep, row, outer axes
  const double *qc;       // dt/dx(i) of this sweep, i = 1..N (prep kernel)
  // Uniform axes (Enzo grids have one cell width per axis): every table
  // entry of the sweep / row / outer axis is the same double, passed here so
  // that no table is read.  uni[]: sweep, row, outer.
  int uni[3];
  double kt[NTAB], kt_row[NTAB], kt_out[NTAB];
  // a uniform sweep axis whose width is a power of two: 1/dx and 1/(2 dx)
  // are exact, so x / dx == x * inv_dx and x / (2 dx) == x * inv_2dx bit for
  // bit (both are the correctly rounded value of the same real number);
  // set by set_dx_inverse
  int dx_pow2;
  double inv_dx, inv_2dx;
  double gamma, eta1, eta2, pmin, dfloor, dt;
  const double *dtp;      // if non-null, dt = *dtp
  int isteep, iflatten, idiff, idual, ipresfree, gravity, ifallback;
  int nact_out;           // active cells along the outer axis (calcdiss nzz)
  const int *face_head;   // [N+2]: records of position i are [head[i], head[i+1])
  const FaceRec *faces;
  double *face_out;
  unsigned long long *counters;  // [0] fallback cells [1] threads that redid a dual-energy
                                 // chain (diagnostic) [2] errors
                                 // [3] fallback needed on a pencil too long for it
  // pencil groups (block / nseg of the main launch) to redo with the
  // per-cell fallback: flag per group, list of flagged groups, its length,
  // and the count of finished fallback workgroups (which clear both)
  int *fb_flag;
  unsigned *fb_list, *fb_count, *fb_done;
  unsigned fb_tag;  // added to list entries (the grid's first group in a batched level)
  int fb_ok;        // a one-pencil workgroup holds the whole sweep axis
  int hll_all;      // RiemannSolver HLL: every cell takes flux_hll's fluxes
                    // (Grid_xEulerSweep.C:245-256), in one-pencil workgroups
  int hllc;         // RiemannSolver HLLC: flux_hllc's fluxes with its per-cell
                    // flux_hll fallback (Grid_xEulerSweep.C:258-269), likewise
  int tp;           // pencils per group (the sweep tile's TP)
  // learned two-shock start: network parameters (ppm_learn.hpp), read by
  // the LEARN instantiation only
  const float *learn;
  // harvest of interface states: records of kHarvestDoubles doubles for
  // one in `harvest_keep` owned interfaces (by a hash of their position),
  // up to harvest_cap records; harvest_n counts them.  Off when null.
  double *harvest;
  unsigned long long *harvest_n;
  long long harvest_cap;
  unsigned harvest_keep;
};

// Kernels that read their SweepParams from memory (the batched level's)
// see generic pointers there, which the compiler addresses with flat
// instructions (counted with the LDS operations too, so that LDS waits also
// wait for them); each global access below casts its pointer through the
// global address space and accesses memory through the global pointer, so
// that it compiles to a global instruction as for the per-grid kernel's
// by-value parameters.  Identity outside the device pass.
#if defined(__HIP_DEVICE_COMPILE__)
#define PPM_GLOBAL __attribute__((address_space(1)))
#else
#define PPM_GLOBAL
#endif
template <class T>
PPM_HD PPM_GLOBAL T *as_global(T *p) {
  return (PPM_GLOBAL T *)p;
}

// Learned start in the sweep tile: the network's inputs of the interfaces
// that pass the gate are listed in LDS, kNetStride floats each (their
// corrections then overwrite the first), and the workgroup evaluates the
// network on the list together (learn_eval).
constexpr int kNetStride = 6;

// harvest record: dls, drs, pls, prs, uls, urs (the two-shock inputs),
// p* (converged), starting pressure, Newton updates, restarted (0/1)
constexpr int kHarvestDoubles = 10;

// ---------------------------------------------------------------------------
// LDS layout: NARR arrays of [TP][LW] doubles.
enum Arr {
  A_D = 0, A_E, A_U, A_V, A_W, A_GE, A_P, A_CS,
  A_ST, A_DQ, A_QL, A_QR, A_Q6, A_COL, A_X, NARR
};
// flux arrays alias arrays that are dead after the interface solve
constexpr int F_DF = A_DQ, F_EF = A_QL, F_UF = A_QR, F_VF = A_Q6, F_WF = A_ST,
              F_GEF = A_P, F_UB = A_X, F_COLF = A_ST, A_EOUT = A_X, A_GE1 = A_DQ,
              A_HOT = A_QR;
// the learned start's acoustic quantities C_r and p_ac (Acoustic) of each
// interface, between the two stages of the Riemann phase, in arrays dead
// then and outside the network's lists (A_P .. A_Q6): A_COL before the
// fallback flags and the colours, A_X before the flux store (each
// work-item reads its own interface's entries before it stores that
// interface's fluxes into F_UB = A_X); C_l is computed again
constexpr int A_AC_R = A_COL, A_AC_P = A_X;

// Fallback workgroups (one pencil, TP = 1) keep two more arrays: the HLL
// density flux for the in-order fallback decision, then the two gas-energy
// source terms sl*bm0 and sr*bp0 of each interface (flux_hll.F:297-308).
constexpr int A_HA = NARR, A_HB = NARR + 1, NARR_FB = NARR + 2;

// Tile shape of the per-grid sweep kernel: 4 pencils x 64 interfaces (256
// work-items).  Longer segments cut the halo recomputation (segments of up to
// 57 owned cells); on the 256^3 workload 4 x 64 beat 8 x 32, 2 x 128 and
// 4 x 96 (3.26 ms against 2.40 per sweep, though its 3 segments of a
// 262-cell pencil take 10% fewer work-items than 5 x 64: pencils that
// straddle wavefronts cost more than the idle lanes saved).
constexpr int kGridTP = 4, kGridTS = 64;

// LW: cells per pencil in LDS.  A tile loads at most NI + 5 cells (two
// halo cells below its first interface and three above its last) and
// build_segments keeps NI + 4 <= TS, so TS + 1 cells hold any tile.
template <int TP, int TS, int UNI = 0>
struct Cfg {
  static constexpr int LW = TS + 1;
  static constexpr int NT = TP * TS;
  static constexpr int SMEM = NARR * TP * LW;  // doubles
};

struct Tile {
  int o, w;          // owned cells [o, o+w)
  int IA, IB, NI;    // interfaces [IA, IB]
  int L0, L1;        // loaded cells [L0, L1]
  int t1_0, t2;      // first pencil (t1) of the block, pencil row t2
  int np;            // pencils in this block
  double dt;         // time step, read once per block (it may live on the device)
  double qcu;        // dt/dx of a uniform sweep axis (qc_of)
};

// characteristic distances of one cell (inteuler.F:201-208)
struct Chars {
  double ch1, ch2, cm, c0, cp;
};
PPM_HD Chars chars_of(double u, double cs, double dt, double dx2i) {
  Chars r;
  r.ch1 = fmx(0.0, dt * (u + cs)) * dx2i;
  r.ch2 = fmx(0.0, -(dt * (u - cs))) * dx2i;
  r.cm = dt * (u - cs) * dx2i;
  r.c0 = dt * (u)*dx2i;
  r.cp = dt * (u + cs) * dx2i;
  return r;
}

// The acoustic quantities of an interface for the learned start:
// C_l = sqrt(gamma p_l rho_l), C_r, and the start p_ac floored at pmin
// (twoshock.F:156-170).
struct Acoustic {
  double cl, cr, ps;
};

struct Regs {
  // interface averages (inteuler.F before the corrections)
  double dla, dra, dl0, dr0;
  double pla, pra, pl0, pr0, plm, prm, plp, prp;
  double ula, ura, ulm, urm, ulp, urp;
  double vls, vrs, wls, wrs, gels, gers;
  // kept for the colour passes and the flux store
  double dls_pre, drs_pre, dls, drs, db, ub;
  double df, ef, uf, vf, wf, gef;
  int ovl, ovr, upl, upr;  // DE overrides; upwind choice (u(i-1)<=0, u(i)>=0)
  Chars ca, cb;            // characteristic distances of cells i-1 and i
  // two-shock inputs kept between the split Riemann phases of the learned
  // start (dls / drs are above), and the interface's slot in the workgroup's
  // list of network evaluations (-1: none)
  double s_pl, s_pr, s_ul, s_ur;
  int ls;
  float lc;  // the network's correction at that slot
  double flat_c;           // flattening at this thread's parabola cell
  // fallback workgroups only: HLL fluxes of the interface (dt/dx absorbed),
  // HLL weights and wave-speed offsets, and whether they replace the
  // two-shock fluxes
  double h_df, h_uf, h_vf, h_wf, h_ef, h_gef, h_sl, h_sr, h_bm0, h_bp0;
  int hll;
  // HLLC workgroups: the HLLC weights of the interface, the gas-energy
  // source terms sl*gesl and sr*gesr, and the cell's gas-energy source
  double c_sl, c_sr, c_ga, c_gb, c_ges;
};

// Tile LDS accessors.  A checking build (PPM_LDS_TRACK, defined by the CPU
// execution of the HIP source in tests/emu) hands out a reference that
// reports every read and write to ppm_lds_note_read / ppm_lds_note_write,
// so that a read of a cell the workgroup has not written is caught; a kernel
// announces its tile array with PPM_LDS_ENTER (nothing in the HIP build).
#ifdef PPM_LDS_TRACK
struct LdsRef {
  double *q;
  operator double() const {
    ppm_lds_note_read(q);
    return *q;
  }
  LdsRef &operator=(double v) {
    ppm_lds_note_write(q);
    *q = v;
    return *this;
  }
  LdsRef &operator=(const LdsRef &o) { return *this = (double)o; }
};
#define PPM_LDS_REF LdsRef
#define PPM_LDS_AT(...) LdsRef{&(__VA_ARGS__)}
#define PPM_LDS_ENTER(arr) ppm_lds_enter((arr), sizeof(arr))
#else
#define PPM_LDS_ENTER(arr) ((void)0)
#define PPM_LDS_REF double &
#define PPM_LDS_AT(...) (__VA_ARGS__)
#endif

template <int TP, int TS, int UNI = 0>
PPM_HD PPM_LDS_REF LDS(double *sm, int arr, int p, int c) {
  PPM_ASSERT(arr >= 0 && arr < NARR && p >= 0 && p < TP && c >= 0 && c < Cfg<TP, TS, UNI>::LW);
  return PPM_LDS_AT(sm[(arr * TP + p) * Cfg<TP, TS, UNI>::LW + c]);
}

// the fallback arrays A_HA / A_HB (one-pencil tiles only)
template <int TP, int TS, int UNI = 0>
PPM_HD PPM_LDS_REF LDSH(double *sm, int arr, int c) {
  static_assert(TP == 1, "fallback arrays exist in one-pencil tiles only");
  PPM_ASSERT((arr == A_HA || arr == A_HB) && c >= 0 && c < Cfg<TP, TS, UNI>::LW);
  return PPM_LDS_AT(sm[arr * Cfg<TP, TS, UNI>::LW + c]);
}

PPM_HD double dtval(const SweepParams &P) { return P.dtp ? *as_global(P.dtp) : P.dt; }

template <class Ptr>
PPM_HD double tabv(Ptr tab, int N, int t, int cell) {
  PPM_ASSERT(cell >= 1 && cell <= N && t >= 0 && t < NTAB);
  return tab[t * N + (cell - 1)];
}

// coefficient t at `cell` of the sweep axis (row / outer axis for TR / TO);
// UNI: the sweep axis has one cell width (kernel instantiated for it)
template <int UNI>
PPM_HD double TS_(const SweepParams &P, int t, int cell) {
  if (UNI) return P.kt[t];
  return tabv(as_global(P.tab), P.N, t, cell);
}
PPM_HD double TR_(const SweepParams &P, int t, int cell, int n) {
  return P.uni[1] ? P.kt_row[t] : tabv(as_global(P.tab_row), n, t, cell);
}
PPM_HD double TO_(const SweepParams &P, int t, int cell, int n) {
  return P.uni[2] ? P.kt_out[t] : tabv(as_global(P.tab_out), n, t, cell);
}
// whether x is a power of two (then 1/x is exact)
inline bool is_pow2(double x) {
  int e;
  return x > 0.0 && std::frexp(x, &e) == 0.5;
}

// SweepParams::dx_pow2 / inv_dx / inv_2dx from uni[0] and kt[T_DX]
inline void set_dx_inverse(SweepParams &S) {
  S.dx_pow2 = S.uni[0] && is_pow2(S.kt[T_DX]) && S.kt[T_SDXP] == 2.0 * S.kt[T_DX] &&
              S.kt[T_SDXM] == 2.0 * S.kt[T_DX];
  S.inv_dx = S.dx_pow2 ? 1.0 / S.kt[T_DX] : 0.0;
  S.inv_2dx = S.dx_pow2 ? 1.0 / S.kt[T_SDXP] : 0.0;
}

// dt/dx of a uniform sweep axis, once per tile
PPM_HD double qc_uniform(const SweepParams &P, double dt) {
  if (!P.uni[0]) return 0.0;
  return P.dx_pow2 ? dt * P.inv_dx : dt / P.kt[T_DX];
}

// dt/dx(i) (flux_twoshock.F:366)
template <int UNI>
PPM_HD double qc_of(const SweepParams &P, const Tile &T, int cell) {
  if (UNI) return T.qcu;
  return as_global(P.qc)[cell - 1];
}

// grid strides of the axes of a sweep (x fastest); after dim, ax_t1, ax_t2
inline void set_strides(SweepParams &S) {
  const int st[3] = {1, S.GD[0], S.GD[0] * S.GD[1]};
  S.st_s = st[S.dim];
  S.st_1 = st[S.ax_t1];
  S.st_2 = st[S.ax_t2];
}

// 3-D grid index of pencil (t1, t2) at sweep cell `cell` (1-based).  Every
// index is below prod(GridDimension) <= 1031^3 < 2^31, so 32-bit arithmetic.
PPM_HD long long gidx(const SweepParams &P, int t1, int t2, int cell) {
#ifdef PPM_EMU_CHECKS
  int c[3];
  c[P.dim] = cell - 1;
  c[P.ax_t1] = t1;
  c[P.ax_t2] = t2;
  PPM_ASSERT(c[0] >= 0 && c[0] < P.GD[0] && c[1] >= 0 && c[1] < P.GD[1] && c[2] >= 0 &&
             c[2] < P.GD[2]);
  PPM_ASSERT((cell - 1) * P.st_s + t1 * P.st_1 + t2 * P.st_2 ==
             (c[2] * P.GD[1] + c[1]) * P.GD[0] + c[0]);
#endif
  return (long long)((cell - 1) * P.st_s + t1 * P.st_1 + t2 * P.st_2);
}

PPM_HD int coord_of(const SweepParams &P, int axis, int t1, int t2) {
  return axis == P.ax_t1 ? t1 : t2;
}

// Workgroups are handed to the 8 XCDs of the chip round-robin (workgroup b
// runs on XCD b % 8, and every XCD has its own L2).  Tiles are renumbered
// so that each XCD works through one contiguous range of them: neighbouring
// segments of a pencil (which share halo cells) and neighbouring pencil
// groups (which share cache lines when t1 is the contiguous axis) then meet
// in the same L2.  A bijection of [0, nblk).
constexpr int kXCD = 8;
PPM_HD int xcd_tile(int b, int nblk) {
  const int q = nblk / kXCD, rem = nblk % kXCD;
  const int x = b % kXCD, k = b / kXCD;
  return x * q + (x < rem ? x : rem) + k;
}

// interface and load ranges of a tile that owns cells [T.o, T.o + T.w)
PPM_HD void tile_ranges(const SweepParams &P, Tile &T) {
  T.IA = T.o - 1 > P.i1 ? T.o - 1 : P.i1;
  T.IB = T.o + T.w + 1 < P.i2 + 1 ? T.o + T.w + 1 : P.i2 + 1;
  T.NI = T.IB - T.IA + 1;
  int lo = T.IA - 3, lo2 = T.o - 1 > 1 ? T.o - 1 : 1;
  T.L0 = lo < lo2 ? lo : lo2;
  int hi = T.IB + 2, hi2 = T.o + T.w < P.N ? T.o + T.w : P.N;
  T.L1 = hi > hi2 ? hi : hi2;
}

// tile `block` (after xcd_tile) of the launch: segment fastest, then the
// pencil group along t1, then t2.  block / nseg is the pencil group.
template <int TP, int TS, int UNI = 0>
PPM_HD Tile make_tile(const SweepParams &P, int block) {
  Tile T;
  int seg = block % P.nseg;
  int r = block / P.nseg;
  int tb = r % P.nblk_t1;
  T.t2 = P.t2_0 + r / P.nblk_t1;
  T.t1_0 = tb * TP;
  T.np = P.n_t1 - T.t1_0 < TP ? P.n_t1 - T.t1_0 : TP;
  T.o = as_global(P.seg_o)[seg];
  T.w = as_global(P.seg_w)[seg];
  tile_ranges(P, T);
  T.dt = dtval(P);
  T.qcu = qc_uniform(P, T.dt);
  return T;
}

// one whole pencil (t1, t2) as a single tile owning cells [1, N]: the tile
// of the fallback workgroups.  Fits when N - 1 <= TS (interfaces and slope
// items) and N <= Cfg::LW (loaded cells).
PPM_HD Tile make_pencil_tile(const SweepParams &P, int t1, int t2) {
  Tile T;
  T.t2 = t2;
  T.t1_0 = t1;
  T.np = 1;
  T.o = 1;
  T.w = P.N;
  tile_ranges(P, T);
  T.dt = dtval(P);
  T.qcu = qc_uniform(P, T.dt);
  return T;
}

// Items k in [0, n) of every pencil of the tile, spread over the workgroup
// and coalesced along the contiguous grid axis: along the items when
// sweeping x (thread tid takes pencil tid / TS and items tid % TS + m TS),
// along the pencils otherwise (consecutive threads, consecutive pencils).
// TP and TS are compile-time constants, so no integer division remains.
template <int TP, int TS, class F>
PPM_HD void per_cell_coalesced(const SweepParams &P, const Tile &T, int tid, int n, F f) {
  if (P.dim == 0) {
    const int p = tid / TS;
    if (p < T.np)
      for (int k = tid % TS; k < n; k += TS) f(p, k);
  } else {
    for (int idx = tid; idx < TP * n; idx += TP * TS) {
      const int p = idx % TP, k = idx / TP;
      if (p < T.np) f(p, k);
    }
  }
}

// ---------------------------------------------------------------------------
// Phase: load the pencil segments into LDS (coalesced along the contiguous
// grid axis: s when sweeping x, t1 otherwise).
template <int TP, int TS, int UNI = 0>
PPM_HD double flat_of(const SweepParams &P, const Tile &T, int p, int cell);

// the thread's parabola cell (same mapping as ph_parab: one item per thread)
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_flat_reg(const SweepParams &P, const Tile &T, int tid, Regs &r) {
  const int n = T.NI + 1;
  r.flat_c = 0.0;
  const int p = tid / TS, k = tid % TS;
  if (P.iflatten == 0 || p >= T.np || k >= n) return;
  r.flat_c = flat_of<TP, TS, UNI>(P, T, p, T.IA - 1 + k);
}

template <int TP, int TS, int UNI = 0>
PPM_HD void ph_load(const SweepParams &P, const Tile &T, double *sm, int tid) {
  const int nc = T.L1 - T.L0 + 1;
  const int vax = (P.dim + 1) % 3, wax = (P.dim + 2) % 3;
  per_cell_coalesced<TP, TS>(P, T, tid, nc, [&](int p, int c) {
    int cell = T.L0 + c;
    long long g = gidx(P, T.t1_0 + p, T.t2, cell);
    const auto A = as_global(P.A);
    LDS<TP, TS, UNI>(sm, A_D, p, c) = A[S_D * P.cells + g];
    LDS<TP, TS, UNI>(sm, A_E, p, c) = A[S_E * P.cells + g];
    LDS<TP, TS, UNI>(sm, A_U, p, c) = A[(S_VX + P.dim) * P.cells + g];
    LDS<TP, TS, UNI>(sm, A_V, p, c) = vax < P.rank ? A[(S_VX + vax) * P.cells + g] : 0.0;
    LDS<TP, TS, UNI>(sm, A_W, p, c) = wax < P.rank ? A[(S_VX + wax) * P.cells + g] : 0.0;
    if (P.idual) LDS<TP, TS, UNI>(sm, A_GE, p, c) = A[S_GE * P.cells + g];
    LDS<TP, TS, UNI>(sm, A_P, p, c) = as_global(P.pres)[g];
  });
}

// Phase: sound speed (inteuler.F:199-200) at cells [IA-1, IB] into A_CS,
// and density second differences (inteuler.F:166-173) at [IA-2, IB+1]
// (A_DQ).  The characteristic distances of inteuler.F:201-208 are rebuilt
// from (u, cs) where they are used (char_* below), bit for bit.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_chars(const SweepParams &P, const Tile &T, double *sm, int tid) {
  const int nch = T.NI + 1;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < nch; k += TS) {
    int cell = T.IA - 1 + k;
    int c = cell - T.L0;
    double pp = LDS<TP, TS, UNI>(sm, A_P, p, c), dd = LDS<TP, TS, UNI>(sm, A_D, p, c);
    double cs = sqrt(P.gamma * pp / dd);
    if (P.ipresfree == 1) cs = kTiny;
    LDS<TP, TS, UNI>(sm, A_CS, p, c) = cs;
  }
  if (P.isteep) {
    const int n2 = T.NI + 3;
    for (int p = tid / TS, k = tid % TS; p < T.np && k < n2; k += TS) {
      int cell = T.IA - 2 + k;
      int c = cell - T.L0;
      double dm = LDS<TP, TS, UNI>(sm, A_D, p, c - 1), d0 = LDS<TP, TS, UNI>(sm, A_D, p, c),
             dp = LDS<TP, TS, UNI>(sm, A_D, p, c + 1);
      double d2d, d2m;
      if (UNI && P.dx_pow2) {  // x / (2 dx) == x * inv_2dx exactly
        d2d = (dp - d0) * P.inv_2dx;
        d2m = (d0 - dm) * P.inv_2dx;
      } else {
        d2d = (dp - d0) / TS_<UNI>(P, T_SDXP, cell);
        d2m = (d0 - dm) / TS_<UNI>(P, T_SDXM, cell);
      }
      d2d = (d2d - d2m) / TS_<UNI>(P, T_QA3, cell);
      LDS<TP, TS, UNI>(sm, A_DQ, p, c) = d2d;
    }
  }
}

// Phase: steepening coefficients (inteuler.F:174-192) at [IA-1, IB].
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_steepen(const SweepParams &P, const Tile &T, double *sm, int tid) {
  if (!P.isteep) return;
  const int n = T.NI + 1;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = T.IA - 1 + k;
    int c = cell - T.L0;
    double dm = LDS<TP, TS, UNI>(sm, A_D, p, c - 1), dp = LDS<TP, TS, UNI>(sm, A_D, p, c + 1);
    double pm = LDS<TP, TS, UNI>(sm, A_P, p, c - 1), pq = LDS<TP, TS, UNI>(sm, A_P, p, c + 1);
    double d2m = LDS<TP, TS, UNI>(sm, A_DQ, p, c - 1), d2p = LDS<TP, TS, UNI>(sm, A_DQ, p, c + 1);
    double qc = fabs(dp - dm) - 0.01 * fmn(fabs(dp), fabs(dm));
    double s1 = (d2m - d2p) * TS_<UNI>(P, T_STNUM, cell) /
                (TS_<UNI>(P, T_STDEN, cell) * (dp - dm + kTiny));
    if (d2p * d2m > 0.0) s1 = 0.0;
    if (qc <= 0.0) s1 = 0.0;
    double s2 = fmx(0.0, fmn(20.0 * (s1 - 0.05), 1.0));
    double qa = fabs(dp - dm) / fmn(dp, dm);
    double qb = fabs(pq - pm) / fmn(pq, pm);
    LDS<TP, TS, UNI>(sm, A_ST, p, c) = (P.gamma * 0.1 * qa >= qb) ? s2 : 0.0;
  }
}

// Phase: monotonised average slopes of LDS array `arr` (intvar.F:82-94) at
// cells [IA-2, IB+1], into A_DQ.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_slope(const SweepParams &P, const Tile &T, double *sm, int tid, int arr) {
  const int n = T.NI + 3;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = T.IA - 2 + k;
    int c = cell - T.L0;
    double qm = LDS<TP, TS, UNI>(sm, arr, p, c - 1), q0 = LDS<TP, TS, UNI>(sm, arr, p, c),
           qp = LDS<TP, TS, UNI>(sm, arr, p, c + 1);
    double qplus = qp - q0, qmnus = q0 - qm, dq;
    if (qplus * qmnus > 0) {
      double qcent = TS_<UNI>(P, T_C1, cell) * qplus + TS_<UNI>(P, T_C2, cell) * qmnus;
      double qvanl = 2.0 * qplus * qmnus / (qmnus + qplus);
      double temp1 = fmn(fmn(fmn(fabs(qcent), fabs(qvanl)), 2.0 * fabs(qmnus)), 2.0 * fabs(qplus));
      dq = temp1 * sgn1(qcent);
    } else {
      dq = 0.0;
    }
    LDS<TP, TS, UNI>(sm, A_DQ, p, c) = dq;
  }
}

// flattening coefficient of `cell` for the block's pencil p: the reference
// passes `flatten` (line 1 of the slice) to intvar for every line
// (inteuler.F:241-289), so it depends only on the slice (outer) index.
template <int TP, int TS, int UNI>
PPM_HD double flat_of(const SweepParams &P, const Tile &T, int p, int cell) {
  int outer = coord_of(P, (P.dim + 2) % 3, T.t1_0 + p, T.t2);
  return as_global(P.flat)[(long long)outer * P.N + (cell - 1)];
}

// Phase: interface values, steepening, monotonisation, flattening, CHECK_LR
// and parabola (intvar.F:98-160) at cells [IA-1, IB] -> A_QL, A_QR, A_Q6.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_parab(const SweepParams &P, const Tile &T, double *sm, int tid, int arr,
                     int steep, const Regs &rg) {
  const int n = T.NI + 1;
  PPM_ASSERT(n <= TS);  // one cell per thread, as ph_flat_reg (build_segments)
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = T.IA - 1 + k;
    int c = cell - T.L0;
    double qm = LDS<TP, TS, UNI>(sm, arr, p, c - 1), q = LDS<TP, TS, UNI>(sm, arr, p, c),
           qp = LDS<TP, TS, UNI>(sm, arr, p, c + 1);
    double dqm = LDS<TP, TS, UNI>(sm, A_DQ, p, c - 1), dq0 = LDS<TP, TS, UNI>(sm, A_DQ, p, c),
           dqp = LDS<TP, TS, UNI>(sm, A_DQ, p, c + 1);
    double ql = TS_<UNI>(P, T_C3, cell) * qm + TS_<UNI>(P, T_C4, cell) * q +
                TS_<UNI>(P, T_C5, cell) * dqm + TS_<UNI>(P, T_C6, cell) * dq0;
    double qr = TS_<UNI>(P, T_C3, cell + 1) * q + TS_<UNI>(P, T_C4, cell + 1) * qp +
                TS_<UNI>(P, T_C5, cell + 1) * dq0 + TS_<UNI>(P, T_C6, cell + 1) * dqp;
    if (steep) {
      double st = LDS<TP, TS, UNI>(sm, A_ST, p, c);
      ql = (1.0 - st) * ql + st * (qm + 0.5 * dqm);
      qr = (1.0 - st) * qr + st * (qp - 0.5 * dqp);
    }
    double temp1 = (qr - q) * (q - ql);
    double temp2 = qr - ql;
    double temp3 = 6.0 * (q - 0.5 * (qr + ql));
    if (temp1 <= 0.0) { ql = q; qr = q; }
    double temp22 = temp2 * temp2;
    double temp23 = temp2 * temp3;
    if (temp22 < temp23) ql = 3.0 * q - 2.0 * qr;
    if (temp22 < -temp23) qr = 3.0 * q - 2.0 * ql;
    if (P.iflatten != 0) {
      double f = rg.flat_c;
      ql = q * f + ql * (1.0 - f);
      qr = q * f + qr * (1.0 - f);
    }
    ql = fmx(fmn(q, qm), ql);
    ql = fmn(fmx(q, qm), ql);
    qr = fmx(fmn(q, qp), qr);
    qr = fmn(fmx(q, qp), qr);
    LDS<TP, TS, UNI>(sm, A_QL, p, c) = ql;
    LDS<TP, TS, UNI>(sm, A_QR, p, c) = qr;
    LDS<TP, TS, UNI>(sm, A_Q6, p, c) = 6.0 * (q - 0.5 * (ql + qr));
  }
}

// domain-of-dependence averages at interface i from cells a = i-1, b = i
// (intvar.F:162-174, inteuler.F:304-317).  kind: 0 = averages (char1/2, c0),
// 1 = plus the cm/cp averages (p and u).
struct IfaceAvg {
  double qla, qra, ql0, qr0, qlm, qrm, qlp, qrp;
};

template <int TP, int TS, int UNI = 0>
PPM_HD IfaceAvg iface_avg(const Chars &A, const Chars &B, double *sm, int p, int ca, int cb,
                          int with_mp) {
  IfaceAvg r;
  double qr_a = LDS<TP, TS, UNI>(sm, A_QR, p, ca), ql_a = LDS<TP, TS, UNI>(sm, A_QL, p, ca);
  double q6_a = LDS<TP, TS, UNI>(sm, A_Q6, p, ca);
  double ql_b = LDS<TP, TS, UNI>(sm, A_QL, p, cb), qr_b = LDS<TP, TS, UNI>(sm, A_QR, p, cb);
  double q6_b = LDS<TP, TS, UNI>(sm, A_Q6, p, cb);
  double dq_a = qr_a - ql_a, dq_b = qr_b - ql_b;
  double ch1 = A.ch1, ch2 = B.ch2;
  double c0a = A.c0, c0b = B.c0;
  r.qla = qr_a - ch1 * (dq_a - (1.0 - kFt * ch1) * q6_a);
  r.qra = ql_b + ch2 * (dq_b + (1.0 - kFt * ch2) * q6_b);
  r.ql0 = qr_a - c0a * (dq_a - (1.0 - kFt * c0a) * q6_a);
  r.qr0 = ql_b - c0b * (dq_b + (1.0 + kFt * c0b) * q6_b);
  if (with_mp) {
    double cma = A.cm, cmb = B.cm;
    double cpa = A.cp, cpb = B.cp;
    r.qlm = qr_a - cma * (dq_a - (1.0 - kFt * cma) * q6_a);
    r.qrm = ql_b - cmb * (dq_b + (1.0 + kFt * cmb) * q6_b);
    r.qlp = qr_a - cpa * (dq_a - (1.0 - kFt * cpa) * q6_a);
    r.qrp = ql_b - cpb * (dq_b + (1.0 + kFt * cpb) * q6_b);
  } else {
    r.qlm = r.qrm = r.qlp = r.qrp = 0.0;
  }
  return r;
}

// The upwind interface states of an advected quantity (inteuler.F:414-434):
// left qla (upl) or ql0, right qra (upr) or qr0, each computed alone with
// its characteristic distance selected first.  qr0's form is qra's with the
// distance -c0 (a - c*y == a + (-c)*y and 1 + k*c == 1 - k*(-c) exactly), so
// the results are bitwise iface_avg's selected ones.
struct IfaceUp {
  double ql, qr;
};

template <int TP, int TS, int UNI = 0>
PPM_HD IfaceUp iface_up(const Chars &A, const Chars &B, double *sm, int p, int ca, int cb,
                        int upl, int upr) {
  const double qr_a = LDS<TP, TS, UNI>(sm, A_QR, p, ca), ql_a = LDS<TP, TS, UNI>(sm, A_QL, p, ca);
  const double q6_a = LDS<TP, TS, UNI>(sm, A_Q6, p, ca);
  const double ql_b = LDS<TP, TS, UNI>(sm, A_QL, p, cb), qr_b = LDS<TP, TS, UNI>(sm, A_QR, p, cb);
  const double q6_b = LDS<TP, TS, UNI>(sm, A_Q6, p, cb);
  const double dq_a = qr_a - ql_a, dq_b = qr_b - ql_b;
  const double cl = upl ? A.ch1 : A.c0, cr = upr ? B.ch2 : -B.c0;
  IfaceUp r;
  r.ql = qr_a - cl * (dq_a - (1.0 - kFt * cl) * q6_a);
  r.qr = ql_b + cr * (dq_b + (1.0 - kFt * cr) * q6_b);
  return r;
}

// Phase: interface averages of variable VAR (0 d, 1 p, 2 u, 3 v, 4 w, 5 ge)
// for the thread's interface; results into Regs.  VAR is a compile-time
// constant, so each variable's phase is its own straight-line code.
template <int VAR, int TP, int TS, int UNI = 0>
PPM_HD void ph_iface(const SweepParams &P, const Tile &T, double *sm, int tid, Regs &r) {
  (void)P;
  int p = tid / TS, q = tid % TS;
  if (p >= T.np || q >= T.NI) return;
  int i = T.IA + q;
  int ca = i - 1 - T.L0, cb = i - T.L0;
  if constexpr (VAR == 0) {
    // the characteristic distances of cells i-1 and i, kept in registers
    // for the other variables, the Riemann solve and the colours
    r.ca = chars_of(LDS<TP, TS, UNI>(sm, A_U, p, ca), LDS<TP, TS, UNI>(sm, A_CS, p, ca), T.dt,
                    TS_<UNI>(P, T_DX2I, i - 1));
    r.cb = chars_of(LDS<TP, TS, UNI>(sm, A_U, p, cb), LDS<TP, TS, UNI>(sm, A_CS, p, cb), T.dt,
                    TS_<UNI>(P, T_DX2I, i));
  }
  if constexpr (VAR >= 3) {
    const IfaceUp a = iface_up<TP, TS, UNI>(r.ca, r.cb, sm, p, ca, cb, r.upl, r.upr);
    if constexpr (VAR == 3) {
      r.vls = a.ql;
      r.vrs = a.qr;
    } else if constexpr (VAR == 4) {
      r.wls = a.ql;
      r.wrs = a.qr;
    } else {
      r.gels = a.ql;
      r.gers = a.qr;
    }
    return;
  }
  IfaceAvg a = iface_avg<TP, TS, UNI>(r.ca, r.cb, sm, p, ca, cb, VAR == 1 || VAR == 2);
  if constexpr (VAR == 0) {
    r.dla = a.qla; r.dra = a.qra; r.dl0 = a.ql0; r.dr0 = a.qr0;
  } else if constexpr (VAR == 1) {
    r.pla = a.qla; r.pra = a.qra; r.pl0 = a.ql0; r.pr0 = a.qr0;
    r.plm = a.qlm; r.prm = a.qrm; r.plp = a.qlp; r.prp = a.qrp;
  } else if constexpr (VAR == 2) {
    r.ula = a.qla; r.ura = a.qra;
    r.ulm = a.qlm; r.urm = a.qrm; r.ulp = a.qlp; r.urp = a.qrp;
    // upwind choice of the advected states (inteuler.F:414-434)
    r.upl = LDS<TP, TS, UNI>(sm, A_U, p, ca) <= 0.0;
    r.upr = LDS<TP, TS, UNI>(sm, A_U, p, cb) >= 0.0;
  }
}

// Newton iteration of twoshock.F:141-254 for one interface from the start
// ps (floored at pmin).  Returns the number of updates taken; 8 means the
// seven updates did not meet the tolerance.
PPM_HD int twoshock_newton(double dls, double drs, double pls, double prs, double uls, double urs,
                           double cl, double cr, double gamma, double pmin, double ps, double &pbar,
                           double &ubar) {
  const double qa = (gamma + 1.0) / (2.0 * gamma);
  const double tolerance = 1.e-14;
  double old_ps = ps;
  double ubl = 0, ubr = 0, dpdul = 0, dpdur = 0;
  int n;
  for (n = 2; n <= 8; n++) {
    double zl = cl * sqrt((1.0 + qa * (ps / pls - 1.0)));
    double zr = cr * sqrt((1.0 + qa * (ps / prs - 1.0)));
    ubl = uls - (ps - pls) / zl;
    ubr = urs + (ps - prs) / zr;
    dpdul = -(4.0 * (zl * (zl * zl)) / dls / (4.0 * (zl * zl) / dls - (gamma + 1.0) * (ps - pls)));
    dpdur = 4.0 * (zr * (zr * zr)) / drs / (4.0 * (zr * zr) / drs - (gamma + 1.0) * (ps - prs));
    ps = ps + (ubr - ubl) * dpdur * dpdul / (dpdur - dpdul);
    if (ps < pmin) ps = pmin;
    double delta_ps = ps - old_ps;
    old_ps = ps;
    if (fabs(delta_ps / ps) < tolerance) break;
  }
  if (ps < pmin) ps = fmn(pls, prs);
  pbar = ps;
  ubar = ubl + (ubr - ubl) * dpdur / (dpdur - dpdul);
  return n - 1;
}

// two-shock Riemann solver for one interface (twoshock.F:141-254): the
// reference's acoustic starting guess, then Newton.  ps0 receives the start;
// returns the Newton updates taken.
PPM_HD int twoshock1(double dls, double drs, double &pls, double &prs, double uls, double urs,
                     double gamma, double pmin, int ipresfree, double &pbar, double &ubar,
                     double &ps0) {
  if (ipresfree == 1) {
    pbar = pmin;
    ubar = 0.5 * (uls + urs);
    pls = pmin;
    prs = pmin;
    ps0 = pmin;
    return 0;
  }
  double cl = sqrt(gamma * pls * dls);
  double cr = sqrt(gamma * prs * drs);
  double ps = (cr * pls + cl * prs + cr * cl * (uls - urs)) / (cr + cl);
  if (ps < pmin) ps = pmin;
  ps0 = ps;
  return twoshock_newton(dls, drs, pls, prs, uls, urs, cl, cr, gamma, pmin, ps, pbar, ubar);
}

// Whether the learned start is used at an interface: the acoustic start
// p_ac (floored at pmin) above the floor, with an acoustic wave strength
// that passes the gate g (|p_ac - p_l| > g p_l or |p_ac - p_r| > g p_r);
// elsewhere p_ac, whose error is second order in that strength, is kept and
// the network is not evaluated.  Where p_ac is at the floor (it predicts
// cavitation) it is kept too.
// The acoustic quantities of an interface (Acoustic), shared by the gate
// and the solver.
PPM_HD Acoustic acoustic_of(double dls, double drs, double pls, double prs, double uls, double urs,
                            double gamma, double pmin) {
  Acoustic a;
  a.cl = sqrt(gamma * pls * dls);
  a.cr = sqrt(gamma * prs * drs);
  a.ps = acoustic_start(dls, drs, pls, prs, uls, urs, a.cl, a.cr);
  if (a.ps < pmin) a.ps = pmin;
  return a;
}

PPM_HD bool learn_gate(const float *w, const Acoustic &A, double pls, double prs, double pmin) {
  const double g = as_global(w)[kOffGate];
  return A.ps > pmin && (fabs(A.ps - pls) > g * pls || fabs(A.ps - prs) > g * prs);
}

// The same solver started from the learned pressure (ppm_learn.hpp): at an
// interface that passed the gate (have_c), the acoustic start p_ac times
// e^c, c being the network's correction; elsewhere p_ac.  When Newton does
// not meet the tolerance from the learned start, p_ac starts it again
// (restarted = 1), so a non-converging interface gets the reference
// method's result.  Returns all updates taken.
// A holds the interface's acoustic_of.
PPM_HD int twoshock_learned(bool have_c, float c, const Acoustic &A, double dls, double drs,
                            double &pls, double &prs, double uls, double urs, double gamma,
                            double pmin, int ipresfree, double &pbar, double &ubar, double &ps0,
                            int &restarted) {
  restarted = 0;
  if (ipresfree == 1) return twoshock1(dls, drs, pls, prs, uls, urs, gamma, pmin, ipresfree, pbar, ubar, ps0);
  const double cl = A.cl, cr = A.cr, ps_ref = A.ps;
  double ps = ps_ref;
  if (have_c) {  // the interface passed the gate (learn_gate)
    ps = ps_ref * (double)mlp_exp(c);
    if (!(ps >= pmin)) ps = pmin;  // NaN too
  }
  ps0 = ps;
  int it = twoshock_newton(dls, drs, pls, prs, uls, urs, cl, cr, gamma, pmin, ps, pbar, ubar);
  if (it > 7 && ps != ps_ref) {
    restarted = 1;
    it += twoshock_newton(dls, drs, pls, prs, uls, urs, cl, cr, gamma, pmin, ps_ref, pbar, ubar);
  }
  return it;
}

// one harvest record (see SweepParams::harvest) for interface i of pencil
// (t1, t2), kept when a hash of its position falls on 0 mod harvest_keep
PPM_HD void harvest_record(const SweepParams &P, int t1, int t2, int i, const double *v) {
  const unsigned h = (unsigned)i * 73856093u ^ (unsigned)t1 * 19349663u ^ (unsigned)t2 * 83492791u ^
                     (unsigned)P.dim * 2654435761u;
  if (P.harvest_keep > 1 && (h % P.harvest_keep) != 0) return;
  const unsigned long long k = PPM_ATOMIC_ADD_U64(P.harvest_n, 1);
  if ((long long)k >= P.harvest_cap) return;
  for (int m = 0; m < kHarvestDoubles; m++) as_global(P.harvest)[k * kHarvestDoubles + m] = v[m];
}

// diffusion coefficient at interface i (calcdiss.F:207-236, 377-404);
// transverse velocities come from the 3-D state: the current slice and the
// next one from A, the previous slice (already swept) from B.
template <int TP, int TS, int UNI = 0>
PPM_HD double diffcoef1(const SweepParams &P, const Tile &T, double *sm, int p, int i) {
  const int rax = (P.dim + 1) % 3, oax = (P.dim + 2) % 3;
  const int t1 = T.t1_0 + p, t2 = T.t2;
  const int row = coord_of(P, rax, t1, t2), outer = coord_of(P, oax, t1, t2);
  const int j = row + 1, k = outer + 1;  // 1-based as in calcdiss
  const int jdim = P.GD[rax], kdim = P.GD[oax];
  double dc = LDS<TP, TS, UNI>(sm, A_U, p, i - 1 - T.L0) - LDS<TP, TS, UNI>(sm, A_U, p, i - T.L0);
  // grid index of sweep cell `cell` at (row, outer) coordinates
  auto at = [&](int rr, int oo, int cell) {
    int c[3];
    c[P.dim] = cell - 1;
    c[rax] = rr;
    c[oax] = oo;
    return ((long long)c[2] * P.GD[1] + c[1]) * P.GD[0] + c[0];
  };
  if (jdim > 1 && 1 < j && j < jdim) {  // nyz > 1 .and. lj1 (dimy > 1 .and. ljy1)
    const auto v = as_global(P.A) + (long long)(S_VX + rax) * P.cells;
    double vdiff1 = (v[at(row - 1, outer, i)] + v[at(row - 1, outer, i - 1)]) -
                    (v[at(row + 1, outer, i)] + v[at(row + 1, outer, i - 1)]);
    dc = dc + (TS_<UNI>(P, T_DIFFX, i) / TR_(P, T_DDEN, j, jdim)) * vdiff1;
  }
  if (P.nact_out > 1 && 1 < k && k < kdim) {  // nzz > 1 .and. lk1 (dimz > 1 .and. lkz1)
    const auto wo = as_global(P.A) + (long long)(S_VX + oax) * P.cells;
    const auto wn = as_global(P.B) + (long long)(S_VX + oax) * P.cells;
    double wdiff1 = (wn[at(row, outer - 1, i)] + wn[at(row, outer - 1, i - 1)]) -
                    (wo[at(row, outer + 1, i)] + wo[at(row, outer + 1, i - 1)]);
    dc = dc + (TS_<UNI>(P, T_DIFFX, i) / TO_(P, T_DDEN, k, kdim)) * wdiff1;
  }
  return 0.1 * fmx(0.0, dc);
}

// HLL fluxes of one interface (flux_hll.F:107-332, i1 = i2 = the cell of
// the fallback): Roe averages, Davis wave speeds, HLL weights, left/right
// fluxes plus diffusion, dt/dx absorbed.  Inputs are the final interface
// states the two-shock solver used.
struct HllIn {
  double dls, drs, pls, prs, uls, urs, vls, vrs, wls, wrs, gels, gers;
  double diffd, diffuu, diffuv, diffuw, diffue, diffuge;  // zero without diffusion
  double gamma, qc;
  int idual;
};

PPM_HD void hll_flux(const HllIn &h, Regs &r) {
  const double gamma1 = h.gamma - 1.0;
  const double gamma1i = 1.0 / gamma1;
  const double sqrtdl = sqrt(h.dls);
  const double sqrtdr = sqrt(h.drs);
  const double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  const double vroe1 = (sqrtdl * h.uls + sqrtdr * h.urs) * isdlpdr;
  const double vroe2 = (sqrtdl * h.vls + sqrtdr * h.vrs) * isdlpdr;
  const double vroe3 = (sqrtdl * h.wls + sqrtdr * h.wrs) * isdlpdr;
  const double v2 = vroe1 * vroe1 + vroe2 * vroe2 + vroe3 * vroe3;
  const double el = gamma1i * h.pls + 0.5 * h.dls * (h.uls * h.uls + h.vls * h.vls + h.wls * h.wls);
  const double er = gamma1i * h.prs + 0.5 * h.drs * (h.urs * h.urs + h.vrs * h.vrs + h.wrs * h.wrs);
  const double hroe = ((el + h.pls) / sqrtdl + (er + h.prs) / sqrtdr) * isdlpdr;
  const double cs = sqrt(gamma1 * fmx((hroe - 0.5 * v2), kTiny));
  const double char1 = vroe1 - cs, char2 = vroe1 + cs;
  const double csl0 = sqrt(h.gamma * h.pls / h.dls);
  const double csr0 = sqrt(h.gamma * h.prs / h.drs);
  const double csl = fmn(h.uls - csl0, char1);
  const double csr = fmx(h.urs + csr0, char2);
  const double bm = fmn(csl, 0.0), bp = fmx(csr, 0.0);
  const double bm0 = h.uls - bm, bp0 = h.urs - bp;
  const double q1 = (bp + bm) / (bp - bm);
  const double sl = 0.5 * (1.0 + q1), sr = 0.5 * (1.0 - q1);
  const double dubl = h.dls * h.uls, dubr = h.drs * h.urs;
  const double dfl = h.dls * bm0, dfr = h.drs * bp0;
  const double ufl = dubl * bm0 + h.pls, ufr = dubr * bp0 + h.prs;
  const double vfl = h.dls * h.vls * bm0, vfr = h.drs * h.vrs * bp0;
  const double wfl = h.dls * h.wls * bm0, wfr = h.drs * h.wrs * bp0;
  const double efl = el * bm0 + h.pls * h.uls, efr = er * bp0 + h.prs * h.urs;
  r.h_df = h.qc * (sl * dfl + sr * dfr + h.diffd);
  r.h_uf = h.qc * (sl * ufl + sr * ufr + h.diffuu);
  r.h_vf = h.qc * (sl * vfl + sr * vfr + h.diffuv);
  r.h_wf = h.qc * (sl * wfl + sr * wfr + h.diffuw);
  r.h_ef = h.qc * (sl * efl + sr * efr + h.diffue);
  r.h_gef = 0.0;
  if (h.idual == 1) {
    const double gefl = bm0 * h.gels * h.dls, gefr = bp0 * h.gers * h.drs;
    r.h_gef = h.qc * (sl * gefl + sr * gefr + h.diffuge);
  }
  r.h_sl = sl;
  r.h_sr = sr;
  r.h_bm0 = bm0;
  r.h_bp0 = bp0;
}

// HLLC fluxes of one interface (flux_hllc.F:104-340): the Roe averages and
// Davis wave speeds of flux_hll, the contact speed cw and pressure cp, the
// HLLC weights (chosen by the sign of cw), left/right fluxes, the contact
// terms sm*cp and sm*cp*cw, dt/dx absorbed; they replace the interface's
// primary fluxes.  Also the weights and gas-energy source terms the cell
// phases need.
PPM_HD void hllc_flux(const HllIn &h, Regs &r) {
  const double gamma1 = h.gamma - 1.0;
  const double gamma1i = 1.0 / gamma1;
  const double sqrtdl = sqrt(h.dls);
  const double sqrtdr = sqrt(h.drs);
  const double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  const double vroe1 = (sqrtdl * h.uls + sqrtdr * h.urs) * isdlpdr;
  const double vroe2 = (sqrtdl * h.vls + sqrtdr * h.vrs) * isdlpdr;
  const double vroe3 = (sqrtdl * h.wls + sqrtdr * h.wrs) * isdlpdr;
  const double v2 = vroe1 * vroe1 + vroe2 * vroe2 + vroe3 * vroe3;
  const double el = gamma1i * h.pls + 0.5 * h.dls * (h.uls * h.uls + h.vls * h.vls + h.wls * h.wls);
  const double er = gamma1i * h.prs + 0.5 * h.drs * (h.urs * h.urs + h.vrs * h.vrs + h.wrs * h.wrs);
  const double hroe = ((el + h.pls) / sqrtdl + (er + h.prs) / sqrtdr) * isdlpdr;
  const double cs = sqrt(gamma1 * fmx((hroe - 0.5 * v2), kTiny));
  const double char1 = vroe1 - cs, char2 = vroe1 + cs;
  const double csl0 = sqrt(h.gamma * h.pls / h.dls);
  const double csr0 = sqrt(h.gamma * h.prs / h.drs);
  const double csl = fmn(h.uls - csl0, char1);
  const double csr = fmx(h.urs + csr0, char2);
  const double bm = fmn(csl, 0.0), bp = fmx(csr, 0.0);
  // contact wave (flux_hllc.F:158-180)
  const double tl = h.pls - (csl - h.uls) * h.dls * h.uls;
  const double tr = h.prs - (csr - h.urs) * h.drs * h.urs;
  const double dl = h.dls * (csl - h.uls);
  const double dr = -h.drs * (csr - h.urs);
  const double q1 = 1.0 / (dl + dr);
  const double cw = (tr - tl) * q1;
  double cp = (dl * tr + dr * tl) * q1;
  double sl, sr, sm;
  if (cw >= 0.0) {
    sl = cw / (cw - bm);
    sr = 0.0;
    sm = -bm / (cw - bm);
  } else {
    sl = 0.0;
    sr = -cw / (bp - cw);
    sm = bp / (bp - cw);
  }
  cp = fmx(cp, 0.0);
  // left/right fluxes (:222-244), HLLC flux plus diffusion and contact
  // terms (:272-302), dt/dx (:321-340)
  const double dubl = h.dls * h.uls, dubr = h.drs * h.urs;
  const double bm0 = h.uls - bm, bp0 = h.urs - bp;
  const double dfl = dubl - bm * h.dls, dfr = dubr - bp * h.drs;
  const double ufl = dubl * bm0 + h.pls, ufr = dubr * bp0 + h.prs;
  const double vfl = h.dls * h.vls * bm0, vfr = h.drs * h.vrs * bp0;
  const double wfl = h.dls * h.wls * bm0, wfr = h.drs * h.wrs * bp0;
  const double efl = el * bm0 + h.pls * h.uls, efr = er * bp0 + h.prs * h.urs;
  r.df = h.qc * (sl * dfl + sr * dfr + h.diffd);
  r.uf = h.qc * ((sl * ufl + sr * ufr + h.diffuu) + sm * cp);
  r.vf = h.qc * (sl * vfl + sr * vfr + h.diffuv);
  r.wf = h.qc * (sl * wfl + sr * wfr + h.diffuw);
  r.ef = h.qc * ((sl * efl + sr * efr + h.diffue) + sm * cp * cw);
  r.gef = 0.0;
  r.c_ga = r.c_gb = 0.0;
  if (h.idual == 1) {
    const double gefl = bm0 * h.gels * h.dls, gefr = bp0 * h.gers * h.drs;
    r.gef = h.qc * (sl * gefl + sr * gefr + h.diffuge);
    r.c_ga = sl * bm0;
    r.c_gb = sr * bp0;
  }
  r.c_sl = sl;
  r.c_sr = sr;
}

// Phase: Eulerian corrections (inteuler.F:321-551), two-shock solve and
// Godunov fluxes (flux_twoshock.F:84-396) for the thread's interface; in
// fallback workgroups (FB) also the HLL fluxes of the interface.
// STAGE 2: the whole phase.  The learned start splits it around the
// network's evaluation: STAGE 0 computes the two-shock inputs into Regs,
// STAGE 1 continues from them (with Regs::ls / lc, the interface's network
// correction).
template <int TP, int TS, int UNI = 0, int FB = 0, int LEARN = 0, int STAGE = 2>
PPM_HD void ph_solve(const SweepParams &P, const Tile &T, double *sm, int tid, Regs &r) {
  int p = tid / TS, q = tid % TS;
  if constexpr (FB != 0 && STAGE != 1) r.hll = 0;
  if (p >= T.np || q >= T.NI) return;
  const int i = T.IA + q;
  const int ca = i - 1 - T.L0, cb = i - T.L0;
  const double gamma = P.gamma, dt = T.dt;
  double pls, prs, uls, urs, dls, drs;
  if constexpr (STAGE != 1) {
  const Chars &CA = r.ca, &CB = r.cb;
  // correction terms (3.7), inteuler.F:321-383
  double cla = sqrt(fmx(gamma * r.pla * r.dla, 0.0));
  double cra = sqrt(fmx(gamma * r.pra * r.dra, 0.0));
  const double rcla = 1.0 / cla;
  double f1 = rcla;
  double betalp = (r.ula - r.ulp) + (r.pla - r.plp) * f1;
  double betalm = (r.ula - r.ulm) - (r.pla - r.plm) * f1;
  double betal0 = (r.pla - r.pl0) * (f1 * f1) + 1.0 / r.dla - 1.0 / r.dl0;
  double grs = 0.0;
  if (P.gravity == 1) {
    const auto gr = as_global(P.gr);
    grs = 0.25 * dt * (gr[gidx(P, T.t1_0 + p, T.t2, i - 1)] + gr[gidx(P, T.t1_0 + p, T.t2, i)]);
    betalp = betalp - grs;
    betalm = betalm - grs;
  }
  f1 = 0.5 * rcla;  // == 0.5/cla exactly (scaling by 2 commutes with rounding)
  betalp = -(betalp * f1);
  betalm = +betalm * f1;
  if (CA.cp <= 0.0) betalp = 0.0;
  if (CA.cm <= 0.0) betalm = 0.0;
  if (CA.c0 <= 0.0) betal0 = 0.0;
  const double rcra = 1.0 / cra;
  f1 = rcra;
  double betarp = (r.ura - r.urp) + (r.pra - r.prp) * f1;
  double betarm = (r.ura - r.urm) - (r.pra - r.prm) * f1;
  double betar0 = (r.pra - r.pr0) * (f1 * f1) + 1.0 / r.dra - 1.0 / r.dr0;
  if (P.gravity == 1) {
    betarp = betarp - grs;
    betarm = betarm - grs;
  }
  f1 = 0.5 * rcra;
  betarp = -(betarp * f1);
  betarm = +betarm * f1;
  if (CB.cp >= 0) betarp = 0.0;
  if (CB.cm >= 0) betarm = 0.0;
  if (CB.c0 >= 0) betar0 = 0.0;
  // corrected states (3.6), inteuler.F:387-396
  pls = r.pla + (betalp + betalm) * (cla * cla);
  prs = r.pra + (betarp + betarm) * (cra * cra);
  uls = r.ula + (betalp - betalm) * cla;
  urs = r.ura + (betarp - betarm) * cra;
  dls = 1.0 / (1.0 / r.dla - (betal0 + betalp + betalm));
  drs = 1.0 / (1.0 / r.dra - (betar0 + betarp + betarm));
  r.dls_pre = dls;
  r.drs_pre = drs;
  r.ovl = r.ovr = 0;
  // dual-energy override (inteuler.F:493-521)
  if (P.idual == 1) {
    const double cma = CA.cm, c0a = CA.c0, cpa = CA.cp;
    const double cmb = CB.cm, c0b = CB.c0, cpb = CB.cp;
    if (gamma * r.pla / r.dla < P.eta2 * (r.ula * r.ula) ||
        fmx(fmx(fabs(cma), fabs(c0a)), fabs(cpa)) < 1.0e-3 || dls / r.dla > 5.0) {
      r.ovl = 1;
      pls = r.pla;
      uls = r.ula;
      dls = r.dla;
    }
    if (gamma * r.pra / r.dra < P.eta2 * (r.ura * r.ura) ||
        fmx(fmx(fabs(cmb), fabs(c0b)), fabs(cpb)) < 1.0e-3 || drs / r.dra > 5.0) {
      r.ovr = 1;
      prs = r.pra;
      urs = r.ura;
      drs = r.dra;
    }
  }
  // floors (inteuler.F:540-545) and pressure-free reset (:556-561)
  pls = fmx(pls, kTiny);
  prs = fmx(prs, kTiny);
  dls = fmx(dls, kTiny);
  drs = fmx(drs, kTiny);
  if (P.ipresfree == 1) {
    dls = r.dla;
    drs = r.dra;
  }
  r.dls = dls;
  r.drs = drs;
  }  // STAGE != 1
  if constexpr (STAGE == 0) {
    r.s_pl = pls;
    r.s_pr = prs;
    r.s_ul = uls;
    r.s_ur = urs;
    if (P.ipresfree != 1) {  // kept in LDS arrays that are dead here (A_AC_*)
      const Acoustic a = acoustic_of(dls, drs, pls, prs, uls, urs, gamma, P.pmin);
      LDS<TP, TS, UNI>(sm, A_AC_R, p, cb) = a.cr;
      LDS<TP, TS, UNI>(sm, A_AC_P, p, cb) = a.ps;
    }
    return;
  }
  if constexpr (STAGE == 1) {
    dls = r.dls;
    drs = r.drs;
    pls = r.s_pl;
    prs = r.s_pr;
    uls = r.s_ul;
    urs = r.s_ur;
  }
  // two-shock solution
  double pbar, ubar, ps0;
  const double pls_in = pls, prs_in = prs;
  int iters, restarted = 0;
  if constexpr (LEARN != 0 && STAGE == 1) {
    Acoustic a;
    a.cl = sqrt(gamma * pls * dls);  // as acoustic_of
    a.cr = LDS<TP, TS, UNI>(sm, A_AC_R, p, cb);
    a.ps = LDS<TP, TS, UNI>(sm, A_AC_P, p, cb);
    iters = twoshock_learned(r.ls >= 0, r.lc, a, dls, drs, pls, prs, uls, urs, gamma, P.pmin,
                             P.ipresfree, pbar, ubar, ps0, restarted);
  }
  else
    iters = twoshock1(dls, drs, pls, prs, uls, urs, gamma, P.pmin, P.ipresfree, pbar, ubar, ps0);
  if (FB == 0 && P.harvest && i >= T.o && i < T.o + T.w) {
    const double v[kHarvestDoubles] = {dls, drs, pls_in, prs_in, uls, urs, pbar, ps0,
                                       (double)iters, (double)restarted};
    harvest_record(P, T.t1_0 + p, T.t2, i, v);
  }
  // Godunov state (flux_twoshock.F:84-218)
  const double qa = (gamma + 1.0) / (2.0 * gamma);
  double sn = sgn1(-ubar);
  double u0, p0, d0;
  if (sn < 0.0) { u0 = uls; p0 = pls; d0 = dls; }
  else { u0 = urs; p0 = prs; d0 = drs; }
  double c0 = sqrt(fmx(gamma * p0 / d0, kTiny));
  double z0 = c0 * d0 * sqrt(fmx(1.0 + qa * (pbar / p0 - 1.0), kTiny));
  double dbar = 1.0 / (1.0 / d0 - (pbar - p0) / fmx(z0 * z0, kTiny));
  double cbar = sqrt(fmx(gamma * pbar / dbar, kTiny));
  double l0, lbar;
  if (pbar < p0) {
    l0 = u0 * sn + c0;
    lbar = sn * ubar + cbar;
  } else {
    l0 = u0 * sn + z0 / d0;
    lbar = l0;
  }
  double frac = l0 - lbar;
  if (frac < kTiny) frac = kTiny;
  frac = (0.0 - lbar) / frac;
  frac = fmn(fmx(frac, 0.0), 1.0);
  double pb = p0 * frac + pbar * (1.0 - frac);
  double db = d0 * frac + dbar * (1.0 - frac);
  double ub = u0 * frac + ubar * (1.0 - frac);
  if (lbar >= 0.0) { pb = pbar; db = dbar; ub = ubar; }
  if (l0 < 0.0) { pb = p0; db = d0; ub = u0; }
  double vb, wb, geb;
  if (ub > 0.0) { vb = r.vls; wb = r.wls; geb = r.gels; }
  else { vb = r.vrs; wb = r.wrs; geb = r.gers; }
  double eb = pb / ((gamma - 1.0) * db) + 0.5 * (ub * ub + vb * vb + wb * wb);
  double upb = pb * ub;
  double dub = ub * db;
  double duub, duvb, duwb, dueb, dugeb = 0.0;
  double dc = 0.0;
  if (P.idiff != 0) {  // flux_twoshock.F:291-322
    dc = diffcoef1<TP, TS, UNI>(P, T, sm, p, i);
    double da = LDS<TP, TS, UNI>(sm, A_D, p, ca), dbb = LDS<TP, TS, UNI>(sm, A_D, p, cb);
    duub = dub * ub + dc * (da * LDS<TP, TS, UNI>(sm, A_U, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_U, p, cb));
    duvb = dub * vb + dc * (da * LDS<TP, TS, UNI>(sm, A_V, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_V, p, cb));
    duwb = dub * wb + dc * (da * LDS<TP, TS, UNI>(sm, A_W, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_W, p, cb));
    dueb = dub * eb + dc * (da * LDS<TP, TS, UNI>(sm, A_E, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_E, p, cb));
    dub = dub + dc * (da - dbb);
    if (P.idual == 1)
      dugeb = dub * geb +
              dc * (da * LDS<TP, TS, UNI>(sm, A_GE, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_GE, p, cb));
  } else {
    duub = dub * ub;
    duvb = dub * vb;
    duwb = dub * wb;
    dueb = dub * eb;
    if (P.idual == 1) dugeb = dub * geb;
  }
  double qc = qc_of<UNI>(P, T, i);
  r.df = qc * dub;
  r.ef = qc * (dueb + upb);
  r.uf = qc * (duub + pb);
  r.vf = qc * duvb;
  r.wf = qc * duwb;
  r.gef = qc * dugeb;
  r.ub = ub;
  r.db = db;
  if constexpr (FB != 0) {
    HllIn h;
    h.dls = dls; h.drs = drs; h.pls = pls; h.prs = prs; h.uls = uls; h.urs = urs;
    h.vls = r.vls; h.vrs = r.vrs; h.wls = r.wls; h.wrs = r.wrs; h.gels = r.gels; h.gers = r.gers;
    h.diffd = h.diffuu = h.diffuv = h.diffuw = h.diffue = h.diffuge = 0.0;
    if (P.idiff != 0) {  // flux_hll.F:166-195
      const double da = LDS<TP, TS, UNI>(sm, A_D, p, ca), dbb = LDS<TP, TS, UNI>(sm, A_D, p, cb);
      h.diffd = dc * (da - dbb);
      h.diffuu = dc * (da * LDS<TP, TS, UNI>(sm, A_U, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_U, p, cb));
      h.diffuv = dc * (da * LDS<TP, TS, UNI>(sm, A_V, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_V, p, cb));
      h.diffuw = dc * (da * LDS<TP, TS, UNI>(sm, A_W, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_W, p, cb));
      h.diffue = dc * (da * LDS<TP, TS, UNI>(sm, A_E, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_E, p, cb));
      if (P.idual == 1)
        h.diffuge = dc * (da * LDS<TP, TS, UNI>(sm, A_GE, p, ca) - dbb * LDS<TP, TS, UNI>(sm, A_GE, p, cb));
    }
    h.gamma = gamma;
    h.qc = qc;
    h.idual = P.idual;
    hll_flux(h, r);
    if (P.hllc) hllc_flux(h, r);
  }
}

// Phase: store the interface fluxes into LDS (indexed by the right cell).
template <int TP, int TS, int UNI = 0, int FB = 0>
PPM_HD void ph_store_flux(const SweepParams &P, const Tile &T, double *sm, int tid,
                          const Regs &r) {
  (void)P;
  int p = tid / TS, q = tid % TS;
  if (p >= T.np || q >= T.NI) return;
  int cb = T.IA + q - T.L0;
  if constexpr (FB != 0) LDSH<TP, TS, UNI>(sm, A_HA, cb) = r.h_df;
  LDS<TP, TS, UNI>(sm, F_DF, p, cb) = r.df;
  LDS<TP, TS, UNI>(sm, F_EF, p, cb) = r.ef;
  LDS<TP, TS, UNI>(sm, F_UF, p, cb) = r.uf;
  LDS<TP, TS, UNI>(sm, F_VF, p, cb) = r.vf;
  LDS<TP, TS, UNI>(sm, F_WF, p, cb) = r.wf;
  LDS<TP, TS, UNI>(sm, F_GEF, p, cb) = r.gef;
  LDS<TP, TS, UNI>(sm, F_UB, p, cb) = r.ub;
  if constexpr (FB != 0) {
    if (P.hllc) {  // the HLLC gas-energy source terms (F_UB is unused here)
      LDS<TP, TS, UNI>(sm, F_UB, p, cb) = r.c_ga;
      LDSH<TP, TS, UNI>(sm, A_HB, cb) = r.c_gb;
    }
  }
}

PPM_HD void upd_range(const SweepParams &P, const Tile &T, int &c0, int &c1);

// HLLC workgroups: each updated cell's gas-energy source
// qc*pcent*(sl*gesl(c) + sr*gesr(c) - sl*gesl(c+1) - sr*gesr(c+1))
// (flux_hllc.F:306-316) into a register, then (after a barrier) into F_UB,
// where the cell phase reads it; A_HB is free again for the fallback.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_hllc_ges(const SweepParams &P, const Tile &T, double *sm, int tid, Regs &r) {
  r.c_ges = 0.0;
  int c0, c1;
  upd_range(P, T, c0, c1);
  if (!P.idual || tid > c1 - c0) return;
  const int cell = c0 + tid, c = cell - T.L0;
  const double pcent = fmx((P.gamma - 1.0) * LDS<TP, TS, UNI>(sm, A_GE, 0, c) * LDS<TP, TS, UNI>(sm, A_D, 0, c), kTiny);
  const double qc = qc_of<UNI>(P, T, cell);
  r.c_ges = qc * pcent *
            (LDS<TP, TS, UNI>(sm, F_UB, 0, c) + LDSH<TP, TS, UNI>(sm, A_HB, c) -
             LDS<TP, TS, UNI>(sm, F_UB, 0, c + 1) - LDSH<TP, TS, UNI>(sm, A_HB, c + 1));
}

template <int TP, int TS, int UNI = 0>
PPM_HD void ph_hllc_ges_store(const SweepParams &P, const Tile &T, double *sm, int tid,
                              const Regs &r) {
  int c0, c1;
  upd_range(P, T, c0, c1);
  if (!P.idual || tid > c1 - c0) return;
  LDS<TP, TS, UNI>(sm, F_UB, 0, c0 + tid - T.L0) = r.c_ges;
}

// updated cells of the tile: [IA, IB-1] within [i1, i2]
PPM_HD void upd_range(const SweepParams &P, const Tile &T, int &c0, int &c1) {
  c0 = T.IA > P.i1 ? T.IA : P.i1;
  c1 = T.IB - 1 < P.i2 ? T.IB - 1 : P.i2;
}

// Phase: negative density / energy check (flux_twoshock.F:401-403).
// Returns the number of failing cells seen by this thread.
template <int TP, int TS, int UNI = 0>
PPM_HD int ph_check(const SweepParams &P, const Tile &T, double *sm, int tid) {
  int c0, c1;
  upd_range(P, T, c0, c1);
  const int n = c1 - c0 + 1;
  int bad = 0;
  if (n <= 0) return 0;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = c0 + k;
    int c = cell - T.L0;
    double d = LDS<TP, TS, UNI>(sm, A_D, p, c), e = LDS<TP, TS, UNI>(sm, A_E, p, c);
    double dfl = LDS<TP, TS, UNI>(sm, F_DF, p, c), dfr = LDS<TP, TS, UNI>(sm, F_DF, p, c + 1);
    if (d + (dfl - dfr) <= 0.0 || e < 0.0) bad++;
  }
  return bad;
}

// Fallback workgroups (one pencil).  The reference checks the cells of the
// pencil in order (flux_twoshock.F:401-456): cell c falls back when
// d + (df(c) - df(c+1)) <= 0 or e < 0, with df(c) already the HLL flux if
// cell c-1 fell back; flux_hll (i1 = i2 = c) then replaces the fluxes of
// interfaces c and c+1 and the gas-energy sources of cells c and c+1, and its
// own check (d + df(c)) - df(c+1) <= 0 (flux_hll.F:349-351) stops the run.
// The HLL flux of an interface does not depend on which call computes it, so
// only the choice of cells is serial: one thread walks the pencil and writes
// each cell's flag (1 = fell back) into A_COL, free until the colour passes.
// Without the fallback (ifallback = 0) a failing cell is an error.
// Returns the number of fallbacks, plus kFbError if any check failed.
constexpr int kFbError = 1 << 24;

template <int TP, int TS, int UNI = 0>
PPM_HD int ph_fb_decide(const SweepParams &P, const Tile &T, double *sm, int tid) {
  static_assert(TP == 1, "one-pencil tiles only");
  if (tid != 0) return 0;
  int nfb = 0, err = 0, prev = 0;
  LDS<TP, TS, UNI>(sm, A_COL, 0, P.i1 - 1 - T.L0) = 0.0;
  LDS<TP, TS, UNI>(sm, A_COL, 0, P.i2 + 1 - T.L0) = 0.0;
  if (P.hll_all) {
    // the HLL solver on the whole pencil: every interface takes the HLL
    // fluxes and flux_hll's own check (flux_hll.F:336-384) is the only one;
    // no cell counts as a fallback
    for (int cell = P.i1; cell <= P.i2; cell++) {
      const int c = cell - T.L0;
      LDS<TP, TS, UNI>(sm, A_COL, 0, c) = 1.0;
      if ((LDS<TP, TS, UNI>(sm, A_D, 0, c) + LDSH<TP, TS, UNI>(sm, A_HA, c)) -
              LDSH<TP, TS, UNI>(sm, A_HA, c + 1) <= 0.0)
        err = 1;
    }
    return err ? kFbError : 0;
  }
  for (int cell = P.i1; cell <= P.i2; cell++) {
    const int c = cell - T.L0;
    const double d = LDS<TP, TS, UNI>(sm, A_D, 0, c), e = LDS<TP, TS, UNI>(sm, A_E, 0, c);
    const double dfl = prev ? LDSH<TP, TS, UNI>(sm, A_HA, c) : LDS<TP, TS, UNI>(sm, F_DF, 0, c);
    const double dfr = LDS<TP, TS, UNI>(sm, F_DF, 0, c + 1);
    int fb = 0;
    // flux_twoshock.F:401-403 checks d + (df(i) - df(i+1)) and e; flux_hllc.F:
    // 344 only (d + df(i)) - df(i+1)
    const bool bad = P.hllc ? (d + dfl) - dfr <= 0.0 : (d + (dfl - dfr) <= 0.0 || e < 0.0);
    if (bad) {
      if (P.ifallback == 0) {
        err = 1;
      } else {
        fb = 1;
        nfb++;
        if ((d + LDSH<TP, TS, UNI>(sm, A_HA, c)) - LDSH<TP, TS, UNI>(sm, A_HA, c + 1) <= 0.0) err = 1;
      }
    }
    LDS<TP, TS, UNI>(sm, A_COL, 0, c) = fb ? 1.0 : 0.0;
    prev = fb;
  }
  return nfb + (err ? kFbError : 0);
}

// Phase: an interface takes its HLL fluxes when the cell on either side fell
// back; every interface stores its gas-energy source terms sl*bm0, sr*bp0
// (flux_hll.F:279-308) into A_HA / A_HB (A_HA's HLL density flux is spent).
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_fb_apply(const SweepParams &P, const Tile &T, double *sm, int tid, Regs &r) {
  (void)P;
  r.hll = 0;
  if (tid >= T.NI) return;
  const int cb = T.IA + tid - T.L0;
  r.hll = (LDS<TP, TS, UNI>(sm, A_COL, 0, cb - 1) != 0.0 || LDS<TP, TS, UNI>(sm, A_COL, 0, cb) != 0.0);
  if (r.hll) {
    LDS<TP, TS, UNI>(sm, F_DF, 0, cb) = r.h_df;
    LDS<TP, TS, UNI>(sm, F_EF, 0, cb) = r.h_ef;
    LDS<TP, TS, UNI>(sm, F_UF, 0, cb) = r.h_uf;
    LDS<TP, TS, UNI>(sm, F_VF, 0, cb) = r.h_vf;
    LDS<TP, TS, UNI>(sm, F_WF, 0, cb) = r.h_wf;
    LDS<TP, TS, UNI>(sm, F_GEF, 0, cb) = r.h_gef;
  }
  LDSH<TP, TS, UNI>(sm, A_HA, cb) = r.h_sl * r.h_bm0;
  LDSH<TP, TS, UNI>(sm, A_HB, cb) = r.h_sr * r.h_bp0;
}

// Phase: conservative update (euler.F:121-217, gravity :248-273) of the
// tile's updated cells, in place in LDS.  Returns error flags.  With fbon (a
// fallback workgroup whose pencil fell back), cells c with a fallback at c
// or c-1 take flux_hll's gas-energy source
// qc*pcent*(sl(c)gesl(c) + sr(c)gesr(c) - sl(c+1)gesl(c+1) - sr(c+1)gesr(c+1)).
// For a cell c+1 after a fallback at c alone, the reference's one-cell call
// evaluates that at c+2 from locals it never set (flux_hll.F:304-307 with
// i2 = c); here they are the HLL terms of interface c+2, as in the oracle.
template <int TP, int TS, int UNI = 0, int FB = 0>
PPM_HD int ph_euler(const SweepParams &P, const Tile &T, double *sm, int tid, bool fbon = false) {
  int c0, c1;
  upd_range(P, T, c0, c1);
  const int n = c1 - c0 + 1;
  int err = 0;
  if (n <= 0) return 0;
  const double dt = T.dt, gamma = P.gamma;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = c0 + k;
    int c = cell - T.L0;
    double d = LDS<TP, TS, UNI>(sm, A_D, p, c), e = LDS<TP, TS, UNI>(sm, A_E, p, c);
    double u = LDS<TP, TS, UNI>(sm, A_U, p, c), v = LDS<TP, TS, UNI>(sm, A_V, p, c),
           w = LDS<TP, TS, UNI>(sm, A_W, p, c);
    double dnu = d + (LDS<TP, TS, UNI>(sm, F_DF, p, c) - LDS<TP, TS, UNI>(sm, F_DF, p, c + 1));
    double dnuinv = 1.0 / dnu;
    if (P.dfloor > 0.0) {
      dnu = fmx(dnu, P.dfloor);
      dnuinv = 1.0 / dnu;
    }
    double uold = u;
    u = (u * d + (LDS<TP, TS, UNI>(sm, F_UF, p, c) - LDS<TP, TS, UNI>(sm, F_UF, p, c + 1))) * dnuinv;
    v = (v * d + (LDS<TP, TS, UNI>(sm, F_VF, p, c) - LDS<TP, TS, UNI>(sm, F_VF, p, c + 1))) * dnuinv;
    w = (w * d + (LDS<TP, TS, UNI>(sm, F_WF, p, c) - LDS<TP, TS, UNI>(sm, F_WF, p, c + 1))) * dnuinv;
    e = fmx(0.1 * e,
            (e * d + (LDS<TP, TS, UNI>(sm, F_EF, p, c) - LDS<TP, TS, UNI>(sm, F_EF, p, c + 1))) * dnuinv);
    if (P.idual == 1) {
      double ge = LDS<TP, TS, UNI>(sm, A_GE, p, c);
      double pcent = fmx((gamma - 1.0) * ge * d, kTiny);
      double qc = qc_of<UNI>(P, T, cell);
      double ges = qc * pcent * (LDS<TP, TS, UNI>(sm, F_UB, p, c) - LDS<TP, TS, UNI>(sm, F_UB, p, c + 1));
      if constexpr (FB != 0) {
        if (P.hllc) ges = LDS<TP, TS, UNI>(sm, F_UB, p, c);  // ph_hllc_ges
        if (fbon &&
            (LDS<TP, TS, UNI>(sm, A_COL, p, c) != 0.0 || LDS<TP, TS, UNI>(sm, A_COL, p, c - 1) != 0.0))
          ges = qc * pcent *
                (LDSH<TP, TS, UNI>(sm, A_HA, c) + LDSH<TP, TS, UNI>(sm, A_HB, c) -
                 LDSH<TP, TS, UNI>(sm, A_HA, c + 1) - LDSH<TP, TS, UNI>(sm, A_HB, c + 1));
      }
      ge = fmx((ge * d + (LDS<TP, TS, UNI>(sm, F_GEF, p, c) - LDS<TP, TS, UNI>(sm, F_GEF, p, c + 1)) + ges) *
                   dnuinv,
               0.5 * ge);
      if (ge < 0.0) err |= 1;
      LDS<TP, TS, UNI>(sm, A_GE, p, c) = ge;
    }
    if (P.gravity == 1) {
      const double gr = as_global(P.gr)[gidx(P, T.t1_0 + p, T.t2, cell)];
      u = u + dt * gr * 0.5 * (d * dnuinv + 1.0);
      e = e + dt * gr * 0.5 * (u + uold * d * dnuinv);
      e = fmx(e, kTiny);
    }
    LDS<TP, TS, UNI>(sm, A_D, p, c) = dnu;
    LDS<TP, TS, UNI>(sm, A_E, p, c) = e;
    LDS<TP, TS, UNI>(sm, A_U, p, c) = u;
    LDS<TP, TS, UNI>(sm, A_V, p, c) = v;
    LDS<TP, TS, UNI>(sm, A_W, p, c) = w;
  }
  return err;
}

// pgas2d_dual range [i1-3, i2+3] (Grid_xEulerSweep.C:300-304) within [1, N]
PPM_HD void pg_range(const SweepParams &P, int &g0, int &g1) {
  g0 = P.i1 - 3 > 1 ? P.i1 - 3 : 1;
  g1 = P.i2 + 3 < P.N ? P.i2 + 3 : P.N;
}

// Phase: speculative synchronised e at cells [o-1, o+w-1] (A_EOUT), with
// ge1 = e - ke (A_GE1) and the flag ge1/e > eta1 (A_HOT) kept for the
// final pass.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_pgas_spec(const SweepParams &P, const Tile &T, double *sm, int tid) {
  int g0, g1;
  pg_range(P, g0, g1);
  int a = T.o - 1 > g0 ? T.o - 1 : g0, b = T.o + T.w - 1 < g1 ? T.o + T.w - 1 : g1;
  int n = b - a + 1;
  if (n <= 0) return;
  for (int p = tid / TS, k = tid % TS; p < T.np && k < n; k += TS) {
    int cell = a + k;
    int c = cell - T.L0;
    const double d = LDS<TP, TS, UNI>(sm, A_D, p, c), e = LDS<TP, TS, UNI>(sm, A_E, p, c);
    const double u = LDS<TP, TS, UNI>(sm, A_U, p, c), v = LDS<TP, TS, UNI>(sm, A_V, p, c),
                 w = LDS<TP, TS, UNI>(sm, A_W, p, c);
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double ge1 = e - ke;
    const bool hot = ge1 / e > P.eta1;
    double ge2 = hot ? ge1 : LDS<TP, TS, UNI>(sm, A_GE, p, c);
    ge2 = fmx(ge2, P.pmin / ((P.gamma - 1.0) * d));
    LDS<TP, TS, UNI>(sm, A_EOUT, p, c) = e - ge1 + ge2;
    LDS<TP, TS, UNI>(sm, A_GE1, p, c) = ge1;
    LDS<TP, TS, UNI>(sm, A_HOT, p, c) = hot ? 1.0 : 0.0;
  }
}

// Phase: dual-energy sync of the owned cells (pgas2d_dual.F:70-118), then
// write every owned cell of the hydro fields to B.  The speculative e is
// exact unless the demax test changed ge in a cell whose e uses ge (never
// with positive energies); those cells are recomputed here and counted.
template <int TP, int TS, int UNI = 0>
PPM_HD int ph_pgas_write(const SweepParams &P, const Tile &T, double *sm, int tid) {
  int g0, g1;
  pg_range(P, g0, g1);
  int chains = 0;
  const int n = T.w;
  const int vax = (P.dim + 1) % 3, wax = (P.dim + 2) % 3;
  per_cell_coalesced<TP, TS>(P, T, tid, n, [&](int p, int k) {
    int cell = T.o + k, c = cell - T.L0;
    double d = LDS<TP, TS, UNI>(sm, A_D, p, c), e = LDS<TP, TS, UNI>(sm, A_E, p, c);
    double u = LDS<TP, TS, UNI>(sm, A_U, p, c), v = LDS<TP, TS, UNI>(sm, A_V, p, c),
           w = LDS<TP, TS, UNI>(sm, A_W, p, c);
    double ge = P.idual ? LDS<TP, TS, UNI>(sm, A_GE, p, c) : 0.0;
    if (P.idual && cell >= g0 && cell <= g1) {
      int im1 = cell - 1 > g0 ? cell - 1 : g0, ip1 = cell + 1 < g1 ? cell + 1 : g1;
      double em = im1 == cell ? d * e
                              : LDS<TP, TS, UNI>(sm, A_D, p, im1 - T.L0) *
                                    LDS<TP, TS, UNI>(sm, A_EOUT, p, im1 - T.L0);
      double ep = LDS<TP, TS, UNI>(sm, A_D, p, ip1 - T.L0) * LDS<TP, TS, UNI>(sm, A_E, p, ip1 - T.L0);
      const double ge1 = LDS<TP, TS, UNI>(sm, A_GE1, p, c);
      const bool hot = LDS<TP, TS, UNI>(sm, A_HOT, p, c) != 0.0;
      double demax = fmx(fmx(d * e, em), ep);
      bool took = ge1 * d / demax > P.eta2;
      double e_new = LDS<TP, TS, UNI>(sm, A_EOUT, p, c);
      if (took) {
        if (!hot) {  // e uses the gas energy just replaced: not speculated
          double ge2 = fmx(ge1, P.pmin / ((P.gamma - 1.0) * d));
          e_new = e - ge1 + ge2;
          chains++;
        }
        ge = ge1;
      }
      e = e_new;
    }
    long long g = gidx(P, T.t1_0 + p, T.t2, cell);
    const auto B = as_global(P.B);
    const auto A = as_global(P.A);
    B[S_D * P.cells + g] = d;
    B[S_E * P.cells + g] = e;
    B[(S_VX + P.dim) * P.cells + g] = u;
    B[(S_VX + vax) * P.cells + g] = vax < P.rank ? v : A[(S_VX + vax) * P.cells + g];
    B[(S_VX + wax) * P.cells + g] = wax < P.rank ? w : A[(S_VX + wax) * P.cells + g];
    if (P.idual) B[S_GE * P.cells + g] = ge;
  });
  return chains;
}

// transverse coordinates used by the face records: the lower and the higher
// of the two transverse axes (Grid_[xyz]EulerSweep.C:319-330 idim, jdim)
PPM_HD void face_coords(const SweepParams &P, int t1, int t2, int &ci, int &cj) {
  int ai = P.dim == 0 ? 1 : 0, aj = P.dim == 2 ? 1 : 2;
  ci = coord_of(P, ai, t1, t2);
  cj = coord_of(P, aj, t1, t2);
}

// Phase: capture of the hydro face fluxes of owned interfaces
// (Grid_xEulerSweep.C:313-376): staging layout per face record is
// [slot][offset] with slots d, e, vx, vy, vz, ge, colours.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_capture(const SweepParams &P, const Tile &T, double *sm, int tid) {
  if (!P.face_head) return;
  int p = tid / TS, q = tid % TS;
  if (p >= T.np || q >= T.NI) return;
  int i = T.IA + q;
  if (i < T.o || i >= T.o + T.w) return;  // owner of interface i = owner of cell i
  int h0 = as_global(P.face_head)[i], h1 = as_global(P.face_head)[i + 1];
  if (h0 == h1) return;
  int ci, cj;
  face_coords(P, T.t1_0 + p, T.t2, ci, cj);
  int cb = i - T.L0;
  const int vax = (P.dim + 1) % 3, wax = (P.dim + 2) % 3;
  for (int h = h0; h < h1; h++) {
    const FaceRec F = as_global(P.faces)[h];
    if (ci < F.lo_i || ci > F.hi_i || cj < F.lo_j || cj > F.hi_j) continue;
    long long off = F.base + (ci - F.lo_i) + (long long)(cj - F.lo_j) * (F.hi_i - F.lo_i + 1);
    as_global(P.face_out)[off + 0 * F.nface] = LDS<TP, TS, UNI>(sm, F_DF, p, cb);
    as_global(P.face_out)[off + 1 * F.nface] = LDS<TP, TS, UNI>(sm, F_EF, p, cb);
    as_global(P.face_out)[off + (2 + P.dim) * F.nface] = LDS<TP, TS, UNI>(sm, F_UF, p, cb);
    as_global(P.face_out)[off + (2 + vax) * F.nface] = LDS<TP, TS, UNI>(sm, F_VF, p, cb);
    as_global(P.face_out)[off + (2 + wax) * F.nface] = LDS<TP, TS, UNI>(sm, F_WF, p, cb);
    if (P.idual) as_global(P.face_out)[off + 5 * F.nface] = LDS<TP, TS, UNI>(sm, F_GEF, p, cb);
  }
}

// Phase: load colour n into A_COL.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_load_col(const SweepParams &P, const Tile &T, double *sm, int tid, int n) {
  const int nc = T.L1 - T.L0 + 1;
  const auto src = as_global(P.A) + (long long)(S_COL0 + n) * P.cells;
  per_cell_coalesced<TP, TS>(P, T, tid, nc, [&](int p, int c) {
    LDS<TP, TS, UNI>(sm, A_COL, p, c) = src[gidx(P, T.t1_0 + p, T.t2, T.L0 + c)];
  });
}

// Phase: colour interface state and flux (inteuler.F:437-450, 493-551;
// flux_twoshock.F:251-259, 375-380; euler.F:173-177 division by dx),
// stored as colf/dx into F_COLF, and captured for owned faces.
template <int TP, int TS, int UNI = 0, int FB = 0>
PPM_HD void ph_col_flux(const SweepParams &P, const Tile &T, double *sm, int tid, int n,
                        const Regs &r) {
  int p = tid / TS, q = tid % TS;
  if (p >= T.np || q >= T.NI) return;
  int i = T.IA + q;
  int ca = i - 1 - T.L0, cb = i - T.L0;
  // each upwind choice selects the operands of one division (the same
  // operations as dividing both ways and selecting the result)
  const IfaceUp a = iface_up<TP, TS, UNI>(r.ca, r.cb, sm, p, ca, cb, r.upl, r.upr);
  double colls = a.ql * r.dls_pre / (r.upl ? r.dla : r.dl0);
  double colrs = a.qr * r.drs_pre / (r.upr ? r.dra : r.dr0);
  if (r.ovl) colls = colls * r.dla / r.dls_pre;
  if (r.ovr) colrs = colrs * r.dra / r.drs_pre;
  colls = fmx(colls, kColorFloor);
  colrs = fmx(colrs, kColorFloor);
  const bool up = r.ub > 0.0;
  double colb = (up ? colls : colrs) * r.db / (up ? r.dls : r.drs);
  double colf = T.dt * r.ub * colb;
  if (FB != 0 && r.hll)  // flux_hll.F:231-236, 272-276, 327-331
    colf = T.dt * (r.h_sl * (r.h_bm0 * colls) + r.h_sr * (r.h_bp0 * colrs));
  else if (FB != 0 && P.hllc)  // flux_hllc.F:252-258, 286-290, 336-340
    colf = T.dt * (r.c_sl * (r.h_bm0 * colls) + r.c_sr * (r.h_bp0 * colrs));
  colf = (UNI && P.dx_pow2) ? colf * P.inv_dx : colf / TS_<UNI>(P, T_DX, i);
  LDS<TP, TS, UNI>(sm, F_COLF, p, cb) = colf;
  if (P.face_head && i >= T.o && i < T.o + T.w) {
    int h0 = as_global(P.face_head)[i], h1 = as_global(P.face_head)[i + 1];
    int ci, cj;
    face_coords(P, T.t1_0 + p, T.t2, ci, cj);
    for (int h = h0; h < h1; h++) {
      const FaceRec F = as_global(P.faces)[h];
      if (ci < F.lo_i || ci > F.hi_i || cj < F.lo_j || cj > F.hi_j) continue;
      long long off = F.base + (ci - F.lo_i) + (long long)(cj - F.lo_j) * (F.hi_i - F.lo_i + 1);
      as_global(P.face_out)[off + (6 + n) * F.nface] = colf;
    }
  }
}

// Phase: colour update (euler.F:178-186) of owned cells and write to B.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_col_update(const SweepParams &P, const Tile &T, double *sm, int tid, int n) {
  const double min_color = 1e-5 * kTiny;
  const auto dst = as_global(P.B) + (long long)(S_COL0 + n) * P.cells;
  per_cell_coalesced<TP, TS>(P, T, tid, T.w, [&](int p, int k) {
    int cell = T.o + k, c = cell - T.L0;
    double col = LDS<TP, TS, UNI>(sm, A_COL, p, c);
    if (cell >= P.i1 && cell <= P.i2) {
      col = col + (LDS<TP, TS, UNI>(sm, F_COLF, p, c) - LDS<TP, TS, UNI>(sm, F_COLF, p, c + 1));
      col = fmx(col, min_color);
    }
    dst[gidx(P, T.t1_0 + p, T.t2, cell)] = col;
  });
}

// flattening coefficient of line 1 of slice `outer` at `cell` (calcdiss.F:
// 193-202 shock flag; iflatten = 1: 409-442, CW84 A1-A2; 2: 447-485, the
// Lagrangean A4-A6; 3: 489-566, the multidimensional A7-A9; the maximum of
// 568-585), reading p, u, and for 2 and 3 also d and e along that line.
struct FlatParams {
  int dim, N, i1, i2;
  int GD[3];
  long long cells;
  int n_outer;
  int iflatten;
  double gamma;
  const double *A, *pres;
  double *flat;
  const double *dx;   // cell widths along the sweep (N)
  double dt;
  const double *dtp;  // device dt if non-null
  double *qc;         // out: dt/dx(i), i = 1..N
};

PPM_HD double flat_value(const FlatParams &F, int outer, int cell) {
  const int rax = (F.dim + 1) % 3, oax = (F.dim + 2) % 3;
  const bool de = F.iflatten >= 2;  // types 2 and 3 read density and energy too
  double pv[9], uv[9], dv[9], ev[9];
  for (int k = -4; k <= 4; k++) {
    int cc = cell + k;
    if (cc < 1 || cc > F.N) {
      pv[k + 4] = 1.0; uv[k + 4] = 0.0; dv[k + 4] = 1.0; ev[k + 4] = 0.0;
      continue;
    }
    int c[3];
    c[F.dim] = cc - 1;
    c[rax] = 0;
    c[oax] = outer;
    long long g = ((long long)c[2] * F.GD[1] + c[1]) * F.GD[0] + c[0];
    PPM_ASSERT(g >= 0 && g < F.cells);
    const auto A = as_global(F.A);
    pv[k + 4] = as_global(F.pres)[g];
    uv[k + 4] = A[(S_VX + F.dim) * F.cells + g];
    dv[k + 4] = de ? A[S_D * F.cells + g] : 1.0;
    ev[k + 4] = de ? A[S_E * F.cells + g] : 0.0;
  }
  auto P = [&](int x) { PPM_ASSERT(x - cell >= -4 && x - cell <= 4); return pv[x - cell + 4]; };
  auto U = [&](int x) { PPM_ASSERT(x - cell >= -4 && x - cell <= 4); return uv[x - cell + 4]; };
  auto D = [&](int x) { PPM_ASSERT(x - cell >= -4 && x - cell <= 4); return dv[x - cell + 4]; };
  auto E = [&](int x) { PPM_ASSERT(x - cell >= -4 && x - cell <= 4); return ev[x - cell + 4]; };
  const double gamma = F.gamma;
  auto ftemp = [&](int x) {
    if (x == F.i1 - 2) x = F.i1 - 1;  // calcdiss.F:424-425, 474-475, 569-570
    if (x == F.i2 + 2) x = F.i2 + 1;
    double qb = fabs(P(x + 1) - P(x - 1)) / fmn(P(x + 1), P(x - 1));
    double wflag = (qb > 0.33 && U(x - 1) > U(x + 1)) ? 1.0 : 0.0;
    if (F.iflatten == 1) {
      double qa;
      if (fabs(P(x + 2) - P(x - 2)) / fmn(P(x + 2), P(x - 2)) < 0.33)
        qa = 1.0;
      else
        qa = (P(x + 1) - P(x - 1)) / (P(x + 2) - P(x - 2));
      double f = fmn(1.0, (qa - 0.75) * 10.0 * wflag);
      return fmx(0.0, f);
    }
    // di = 1/d at x +- 2, and Z (calcdiss.F:458-462, 530-534)
    const double dim2 = 1.0 / D(x - 2), dip2 = 1.0 / D(x + 2);
    const double Z = sqrt((fmx(P(x + 2), P(x - 2)) + 0.5 * (P(x + 2) + P(x - 2)) * (gamma - 1.0)) /
                          fmx(dip2, dim2));
    if (F.iflatten == 2) {
      const int is = x + (int)copysign(2.0, P(x + 1) - P(x - 1));
      const double omega =
          fmx(0.0, 0.75 * (10.0 - (P(x + 1) - P(x - 1)) / (P(x + 2) - P(x - 2))));
      const double kt = (Z + sqrt(gamma * P(is) * D(is))) / Z;
      const double kappa = fmx(0.0, (kt - 2.0) / (kt + 0.01));
      return fmn(wflag * omega, kappa);
    }
    const double dp1 = P(x + 1) - P(x - 1), dp2 = P(x + 2) - P(x - 2);
    const double de1 = E(x + 1) - E(x - 1), de2 = E(x + 2) - E(x - 2);
    double dpp = 0.0, dee = 0.0;
    if (dp2 != 0.0) dpp = dp1 / dp2;
    if (de2 != 0.0) dee = de1 / de2;
    const double omega_tilde = fmx(dpp, dee);
    const int ism = x + (int)copysign(2.0, dp1), isp = x - (int)copysign(2.0, dp1);
    double s = -copysign(1.0, dp1);
    if (dp1 == 0.0) s = 0.0;
    const double sigma_tilde = wflag * fabs(dp2) / fmn(P(x + 2), P(x - 2));
    const double sigma = fmx(0.0, (sigma_tilde - 0.5) / (sigma_tilde + 1.0));
    const double omega = fmx(0.0, 10.0 * (omega_tilde - 0.75));
    const double ZE = s * Z / D(ism) + U(ism) + kTiny;
    const double cj2s = sqrt(gamma * P(isp) / D(isp));
    const double kt = fabs((ZE - U(isp) + s * cj2s) / ZE);
    const double kappa = fmx(0.0, (kt - 2.0) / (kt + 0.01));
    return fmn(fmn(kappa, wflag * omega), wflag * sigma);
  };
  if (P(cell + 1) - P(cell - 1) < 0.0) return fmx(ftemp(cell), ftemp(cell + 1));
  return fmx(ftemp(cell), ftemp(cell - 1));
}

// ---------------------------------------------------------------------------
// The whole tile: phases in order; ex.run(f) runs f(tid, regs) on every
// thread followed by a workgroup barrier; ex.count(f) also returns the number
// of threads whose result is nonzero (__syncthreads_count), ex.first(f) the
// result of thread 0.
struct TileCounts {
  int bad, chains, err, nfb;
};

// ---------------------------------------------------------------------------
// Learned start inside the tile.  Between the two stages of the Riemann
// phase, the LDS arrays A_P .. A_Q6 are free (the reconstruction is done,
// the flux store has not begun); as floats they hold the list of network
// inputs of the interfaces that pass the gate ([NT][kNetStride]) and the
// hidden layers of a chunk of CH of them ([CH][32] twice).  The list's
// length is an int in the first entry of A_COL (pencil 0, no interface's
// entry), cleared in the tile's first phase.  The gate is tested in the
// first stage's phase; the workgroup then evaluates the network on the list
// together, one (interface, hidden unit) pair per work-item, so that the
// network costs registers nowhere and work only where the gate passes.
template <int TP, int TS, int UNI>
struct NetLds {
  static constexpr int NT = TP * TS;
  static constexpr int FLOATS = 2 * (A_Q6 - A_P + 1) * TP * Cfg<TP, TS, UNI>::LW;
  // list, two hidden layers of a chunk of ch interfaces, and (w2) a copy of
  // the transposed second-layer weights
  static constexpr int need(int ch, bool w2) {
    return kNetStride * NT + 2 * kMlpHidden * ch + (w2 ? kMlpHidden * kMlpHidden : 0);
  }
  static constexpr bool fits(int ch, bool w2) { return need(ch, w2) <= FLOATS; }
  static constexpr bool W2 = fits(16, true);  // chunks of at least 16 with the weights in LDS
  static constexpr int CH = fits(32, W2) ? 32 : fits(16, W2) ? 16 : fits(8, W2) ? 8 : 4;
  static_assert(fits(CH, W2), "the learned start's lists fit the free tile arrays");
  static constexpr bool outside(int a) { return a < A_P || a > A_Q6; }
  static_assert(outside(A_AC_R) && outside(A_AC_P),
                "the acoustic quantities live outside the lists");
  PPM_HD static float *base(double *sm) {
    return reinterpret_cast<float *>(sm + (size_t)A_P * TP * Cfg<TP, TS, UNI>::LW);
  }
  PPM_HD static int *count(double *sm) {
    return reinterpret_cast<int *>(sm + (size_t)A_COL * TP * Cfg<TP, TS, UNI>::LW);
  }
};

// The gate at the thread's interface (learn_gate), after the first stage of
// the Riemann phase in the same phase: a passing interface appends its
// network inputs to the list and keeps its slot.
template <int TP, int TS, int UNI = 0>
PPM_HD void ph_learn_gate(const SweepParams &P, const Tile &T, double *sm, int tid, Regs &r) {
  r.ls = -1;
  r.lc = 0.0f;
  const int p = tid / TS, q = tid % TS;
  if (p >= T.np || q >= T.NI || P.ipresfree == 1) return;
  PPM_ASSERT(T.IA - T.L0 > 0);  // no interface's A_AC_R entry is the list's length
  Acoustic a;
  a.ps = LDS<TP, TS, UNI>(sm, A_AC_P, p, T.IA + q - T.L0);
  if (!learn_gate(P.learn, a, r.s_pl, r.s_pr, P.pmin)) return;
  const int k = atomicAdd(NetLds<TP, TS, UNI>::count(sm), 1);
  r.ls = k;
  float x[kMlpIn];
  mlp_features(r.dls, r.drs, r.s_pl, r.s_pr, r.s_ul, r.s_ur, P.gamma, x);
  float *in = NetLds<TP, TS, UNI>::base(sm) + kNetStride * k;
  for (int m = 0; m < kMlpIn; m++) in[m] = x[m];
}

// The network (mlp_eval's arithmetic, operation for operation) on the m
// listed interfaces, chunk by chunk; each correction replaces its inputs'
// first float, where the interface's work-item reads it.
template <int TP, int TS, int UNI, class X>
PPM_HD void learn_eval(const SweepParams &P, double *sm, int m, X &ex) {
  using L = NetLds<TP, TS, UNI>;
  constexpr int H = kMlpHidden;
  float *list = L::base(sm), *h1 = list + kNetStride * L::NT, *h2 = h1 + L::CH * H;
  float *w2s = h2 + L::CH * H;  // the transposed W2 in LDS when L::W2
  const auto w = as_global(P.learn);
  const float *w2 = L::W2 ? w2s : P.learn + kOffW2T;
  for (int c0 = 0; c0 < m; c0 += L::CH) {
    const int mc = m - c0 < L::CH ? m - c0 : L::CH;
    ex.run([&](int tid, Regs &r) {  // hidden layer 1 (and the W2 copy, once)
      (void)r;
      if (L::W2 && c0 == 0)
        for (int idx = tid; idx < H * H; idx += L::NT) w2s[idx] = w[kOffW2T + idx];
      for (int idx = tid; idx < mc * H; idx += L::NT) {
        const int e = idx / H, j = idx % H;
        const float *x = list + kNetStride * (c0 + e);
        float a = w[kOffB1 + j];
        for (int k = 0; k < kMlpIn; k++)
          a += w[kOffW1T + k * H + j] * ((x[k] - w[kOffMean + k]) * w[kOffInvStd + k]);
        h1[idx] = mlp_tanh(a);
      }
    });
    ex.run([&](int tid, Regs &r) {  // hidden layer 2 times the output weights
      (void)r;
      for (int idx = tid; idx < mc * H; idx += L::NT) {
        const int e = idx / H, j = idx % H;
        float a = w[kOffB2 + j];
        // a few weights in flight at a time: the interface states stay live
        // in registers across the network
#pragma unroll 4
        for (int k = 0; k < H; k++) a += w2[k * H + j] * h1[e * H + k];
        h2[idx] = w[kOffW3 + j] * mlp_tanh(a);
      }
    });
    ex.run([&](int tid, Regs &r) {  // the output, summed in mlp_eval's order
      (void)r;
      for (int e = tid; e < mc; e += L::NT) {
        float y = w[kOffB3];
#pragma unroll 4
        for (int j = 0; j < H; j++) y += h2[e * H + j];
        list[kNetStride * (c0 + e)] = y;
      }
    });
  }
  // every listed interface takes its correction (before the second stage,
  // whose flux store overwrites the list)
  ex.run([&](int tid, Regs &r) {
    (void)tid;
    if (r.ls >= 0) r.lc = list[kNetStride * r.ls];
  });
}

// Reconstruction of variable V (0 d, 1 p, 2 u, 3 v, 4 w, 5 ge): slopes,
// parabolae (steepened for the density), interface averages.
template <int V, int TP, int TS, int UNI, class X>
PPM_HD void sweep_var(const SweepParams &P, const Tile &T, double *sm, X &ex) {
  constexpr int arr = V == 0 ? A_D : V == 1 ? A_P : V == 2 ? A_U : V == 3 ? A_V : V == 4 ? A_W : A_GE;
  const int steep = (V == 0 && P.isteep) ? 1 : 0;
  ex.run([&](int tid, Regs &r) { (void)r; ph_slope<TP, TS, UNI>(P, T, sm, tid, arr); });
  ex.run([&](int tid, Regs &r) { ph_parab<TP, TS, UNI>(P, T, sm, tid, arr, steep, r); });
  ex.run([&](int tid, Regs &r) { ph_iface<V, TP, TS, UNI>(P, T, sm, tid, r); });
}

template <int TP, int TS, int UNI, int FB, class X, int LEARN = 0>
PPM_HD TileCounts sweep_body(const SweepParams &P, const Tile &T, double *sm, X &ex) {
  TileCounts tc;
  tc.nfb = 0;
  ex.run([&](int tid, Regs &r) {
    ph_flat_reg<TP, TS, UNI>(P, T, tid, r);
    ph_load<TP, TS, UNI>(P, T, sm, tid);
    if constexpr (LEARN != 0)
      if (tid == 0) *NetLds<TP, TS, UNI>::count(sm) = 0;
  });
  ex.run([&](int tid, Regs &r) { (void)r; ph_chars<TP, TS, UNI>(P, T, sm, tid); });
  ex.run([&](int tid, Regs &r) { (void)r; ph_steepen<TP, TS, UNI>(P, T, sm, tid); });
  sweep_var<0, TP, TS, UNI>(P, T, sm, ex);
  sweep_var<1, TP, TS, UNI>(P, T, sm, ex);
  sweep_var<2, TP, TS, UNI>(P, T, sm, ex);
  sweep_var<3, TP, TS, UNI>(P, T, sm, ex);
  sweep_var<4, TP, TS, UNI>(P, T, sm, ex);
  if (P.idual) sweep_var<5, TP, TS, UNI>(P, T, sm, ex);
  if constexpr (LEARN != 0) {
    // the Riemann phase in two stages around the network (NetLds)
    ex.run([&](int tid, Regs &r) {
      ph_solve<TP, TS, UNI, FB, LEARN, 0>(P, T, sm, tid, r);
      ph_learn_gate<TP, TS, UNI>(P, T, sm, tid, r);
    });
    const int m = *NetLds<TP, TS, UNI>::count(sm);  // the same for every work-item
    if (m > 0) learn_eval<TP, TS, UNI>(P, sm, m, ex);
    ex.run([&](int tid, Regs &r) {
      ph_solve<TP, TS, UNI, FB, LEARN, 1>(P, T, sm, tid, r);
      ph_store_flux<TP, TS, UNI, FB>(P, T, sm, tid, r);
    });
  } else {
    ex.run([&](int tid, Regs &r) {
      ph_solve<TP, TS, UNI, FB, LEARN>(P, T, sm, tid, r);
      ph_store_flux<TP, TS, UNI, FB>(P, T, sm, tid, r);
    });
  }
  if constexpr (FB != 0) {
    if (P.hllc) {
      ex.run([&](int tid, Regs &r) { ph_hllc_ges<TP, TS, UNI>(P, T, sm, tid, r); });
      ex.run([&](int tid, Regs &r) { ph_hllc_ges_store<TP, TS, UNI>(P, T, sm, tid, r); });
    }
  }
  tc.bad = ex.count([&](int tid, Regs &r) { (void)r; return ph_check<TP, TS, UNI>(P, T, sm, tid); });
  bool fbon = false;
  int fberr = 0;
  if constexpr (FB != 0) {
    if (tc.bad || P.hll_all || P.hllc) {  // block-uniform
    const int dec = ex.first([&](int tid, Regs &r) {
      (void)r;
      return ph_fb_decide<TP, TS, UNI>(P, T, sm, tid);
    });
    tc.nfb = dec % kFbError;
    fberr = dec / kFbError;
    fbon = tc.nfb > 0 || P.hll_all;
    if (fbon) ex.run([&](int tid, Regs &r) { ph_fb_apply<TP, TS, UNI>(P, T, sm, tid, r); });
    }
  }
  tc.err = ex.count([&](int tid, Regs &r) {
    (void)r;
    ph_capture<TP, TS, UNI>(P, T, sm, tid);
    return ph_euler<TP, TS, UNI, FB>(P, T, sm, tid, fbon);
  });
  tc.err += fberr;
  if (P.idual) ex.run([&](int tid, Regs &r) { (void)r; ph_pgas_spec<TP, TS, UNI>(P, T, sm, tid); });
  tc.chains =
      ex.count([&](int tid, Regs &r) { (void)r; return ph_pgas_write<TP, TS, UNI>(P, T, sm, tid); });
  for (int n = 0; n < P.ncol; n++) {
    ex.run([&](int tid, Regs &r) { (void)r; ph_load_col<TP, TS, UNI>(P, T, sm, tid, n); });
    ex.run([&](int tid, Regs &r) { (void)r; ph_slope<TP, TS, UNI>(P, T, sm, tid, A_COL); });
    ex.run([&](int tid, Regs &r) { ph_parab<TP, TS, UNI>(P, T, sm, tid, A_COL, 0, r); });
    ex.run([&](int tid, Regs &r) { ph_col_flux<TP, TS, UNI, FB>(P, T, sm, tid, n, r); });
    ex.run([&](int tid, Regs &r) { (void)r; ph_col_update<TP, TS, UNI>(P, T, sm, tid, n); });
  }
  return tc;
}

// A tile of the main launch.  bad > 0 means some cell of the tile needs the
// fallback (or, without it, is an error); its pencils are then redone by a
// fallback workgroup, which also settles the errors of those pencils.
template <int TP, int TS, int UNI, class X, int LEARN = 0>
PPM_HD TileCounts sweep_tile(const SweepParams &P, int block, double *sm, X &ex) {
  return sweep_body<TP, TS, UNI, 0, X, LEARN>(P, make_tile<TP, TS, UNI>(P, block), sm, ex);
}

// One whole pencil in a one-pencil workgroup of TS threads, with the
// in-order per-cell fallback; sm holds NARR_FB arrays of Cfg<1, TS>::LW doubles.
template <int TS, int UNI, class X, int LEARN = 0>
PPM_HD TileCounts sweep_pencil_fb(const SweepParams &P, int t1, int t2, double *sm, X &ex) {
  return sweep_body<1, TS, UNI, 1, X, LEARN>(P, make_pencil_tile(P, t1, t2), sm, ex);
}

// whether a one-pencil tile of TS threads holds the whole sweep axis
template <int TS>
PPM_HD bool pencil_fits(const SweepParams &P) {
  Tile T;
  T.o = 1;
  T.w = P.N;
  tile_ranges(P, T);
  return T.NI + 4 <= TS && T.L1 - T.L0 + 1 <= Cfg<1, TS>::LW && T.L0 >= 1 && T.L1 <= P.N;
}

}  // namespace ppm
