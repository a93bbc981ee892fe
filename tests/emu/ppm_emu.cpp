// Test infrastructure: CPU block emulator of the HIP sweep.
//
// Runs the product's tile code (enzo-dev_amd/csrc/ppm_sweep.hpp, the same
// source the HIP kernel is built from) one workgroup at a time on the CPU,
// each phase for every thread id in turn, with every LDS and global access
// bounds-checked (PPM_EMU_CHECKS).  Host orchestration mirrors
// hipPPMSolvePPM_DE: device-slot layout, coefficient tables, segments,
// flattening rows, stale pressure, rotated sweeps with the A/B swap and the
// face-flux staging.  Used by tests to check tile indexing and arithmetic
// against the oracle before (and independently of) a GPU run.
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/enzo_ppm.h"
#include "../../enzo-dev_amd/csrc/ppm_geometry.hpp"
#include "../../enzo-dev_amd/csrc/ppm_sweep.hpp"

using namespace ppm;

namespace {
constexpr int kTP = kGridTP, kTS = kGridTS;  // the device kernel's tile shape
constexpr int kFbTS = 1024;  // one-pencil fallback workgroups

template <int NT>
struct HostExecT {
  Regs r[NT];
  template <class F>
  void run(F f) {
    for (int t = 0; t < NT; t++) f(t, r[t]);
  }
  template <class F>
  int count(F f) {  // as __syncthreads_count: threads with a nonzero result
    int s = 0;
    for (int t = 0; t < NT; t++) s += f(t, r[t]) != 0;
    return s;
  }
  template <class F>
  int first(F f) {  // thread 0's result
    int v = 0;
    for (int t = 0; t < NT; t++) {
      int x = f(t, r[t]);
      if (t == 0) v = x;
    }
    return v;
  }
};
using HostExec = HostExecT<kTP * kTS>;
}  // namespace

extern "C" int emu_solve_ppm_de(const enzo_ppm_grid *g, const enzo_ppm_globals *gl,
                                int64_t CycleNumber, int64_t NumberOfSubgrids,
                                enzo_fluxes **SubgridFluxes, double **CellWidthTemp,
                                const int64_t *GridGlobalStart, int64_t GravityOn,
                                int64_t NumberOfColours, const int64_t *colnum,
                                double MinimumSupportEnergyCoefficient, int64_t *counters_out) {
  const int rank = (int)g->GridRank;
  int GD[3];
  long long cells = 1;
  for (int d = 0; d < 3; d++) { GD[d] = (int)g->GridDimension[d]; cells *= GD[d]; }
  const int ncol = (int)NumberOfColours;
  const int nslot = S_COL0 + ncol;
  std::vector<long long> fnum(nslot, -1);
  fnum[S_D] = g->DensNum; fnum[S_E] = g->TENum; fnum[S_VX] = g->Vel1Num;
  if (rank > 1) fnum[S_VY] = g->Vel2Num;
  if (rank > 2) fnum[S_VZ] = g->Vel3Num;
  if (gl->DualEnergyFormalism) fnum[S_GE] = g->GENum;
  for (int n = 0; n < ncol; n++) fnum[S_COL0 + n] = colnum[n];
  std::vector<double> A((size_t)nslot * cells, 0.0), B((size_t)nslot * cells, 0.0);
  for (int s = 0; s < nslot; s++)
    if (fnum[s] >= 0) memcpy(&A[(size_t)s * cells], g->BaryonField[fnum[s]], sizeof(double) * cells);
  // stale pressure (Grid_ComputePressure.C:96-166)
  std::vector<double> pres(cells);
  for (long long i = 0; i < cells; i++) {
    double dd = A[S_D * cells + i], ge;
    if (gl->DualEnergyFormalism == 0) {
      double v1 = A[S_VX * cells + i], v2 = rank > 1 ? A[S_VY * cells + i] : 0.0,
             v3 = rank > 2 ? A[S_VZ * cells + i] : 0.0;
      ge = A[S_E * cells + i] - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
    } else {
      ge = A[S_GE * cells + i];
    }
    double p = (gl->Gamma - 1.0) * dd * ge;
    if (p < gl->tiny_number) p = gl->tiny_number;
    double mp = MinimumSupportEnergyCoefficient * (gl->Gamma - 1.0) * dd * dd;
    if (p < mp) p = mp;
    pres[i] = p;
  }
  std::vector<double> tabs[3];
  std::vector<int> so[3], sw[3];
  int uni[3];
  double kt[3][NTAB];
  const bool force_tables = getenv("PPM_EMU_TABLES") != nullptr;
  for (int d = 0; d < 3; d++) {
    build_tables(CellWidthTemp[d], GD[d], tabs[d]);
    uni[d] = (!force_tables && uniform_axis(CellWidthTemp[d], GD[d], tabs[d], kt[d])) ? 1 : 0;
    if (d < rank && GD[d] > 1)
      if (!build_segments<kTP, kTS>(GD[d], (int)g->GridStartIndex[d] + 1,
                                    (int)g->GridEndIndex[d] + 1, so[d], sw[d]))
        return ENZO_PPM_ERR_ARGUMENT;
  }
  std::vector<double> accel;
  if (GravityOn) {
    accel.assign((size_t)3 * cells, 0.0);
    for (int d = 0; d < rank; d++)
      memcpy(&accel[(size_t)d * cells], g->AccelerationField[d], sizeof(double) * cells);
  }
  unsigned long long counters[4] = {0, 0, 0, 0};
  std::vector<double> sm(Cfg<kTP, kTS>::SMEM);
  std::vector<double> smfb((size_t)NARR_FB * Cfg<1, kFbTS>::LW);
  HostExec *ex = new HostExec();
  HostExecT<kFbTS> *exfb = new HostExecT<kFbTS>();
  const long long nact[3] = {g->GridEndIndex[0] - g->GridStartIndex[0] + 1,
                             g->GridEndIndex[1] - g->GridStartIndex[1] + 1,
                             g->GridEndIndex[2] - g->GridStartIndex[2] + 1};
  const int ixyz = (int)(CycleNumber % rank);
  for (int nn = ixyz; nn < ixyz + rank; nn++) {
    const int dim = nn % rank;
    if (nact[dim] <= 1) continue;
    const int N = GD[dim], i1 = (int)g->GridStartIndex[dim] + 1, i2 = (int)g->GridEndIndex[dim] + 1;
    const int rax = (dim + 1) % 3, oax = (dim + 2) % 3;
    std::vector<double> flat((size_t)GD[oax] * N, 0.0), qc(N);
    for (int i = 0; i < N; i++) qc[i] = g->dtFixed / CellWidthTemp[dim][i];
    if (gl->PPMFlatteningParameter) {
      FlatParams F;
      memset(&F, 0, sizeof(F));
      F.dim = dim; F.N = N; F.i1 = i1; F.i2 = i2;
      for (int d = 0; d < 3; d++) F.GD[d] = GD[d];
      F.cells = cells; F.n_outer = GD[oax]; F.A = A.data(); F.pres = pres.data(); F.flat = flat.data();
      F.iflatten = (int)gl->PPMFlatteningParameter; F.gamma = gl->Gamma;
      for (int outer = 0; outer < GD[oax]; outer++)
        for (int cell = i1 - 1; cell <= i2 + 1; cell++)
          flat[(size_t)outer * N + cell - 1] = flat_value(F, outer, cell);
    }
    // face records (same rules as hipPPMSetSubgridFluxes)
    std::vector<FaceRec> recs;
    std::vector<int> rsub, rside;
    long long total = 0;
    const int ai = dim == 0 ? 1 : 0, aj = dim == 2 ? 1 : 2;
    for (int n = 0; n < NumberOfSubgrids; n++) {
      enzo_fluxes *Fx = SubgridFluxes[n];
      for (int side = 0; side < 2; side++) {
        FaceRec R;
        memset(&R, 0, sizeof(R));
        long long face = side == 0 ? Fx->LeftFluxStartGlobalIndex[dim][dim] - GridGlobalStart[dim]
                                   : Fx->RightFluxStartGlobalIndex[dim][dim] - GridGlobalStart[dim] + 1;
        R.pos = (int)face + 1;
        R.lo_i = (int)(Fx->RightFluxStartGlobalIndex[dim][ai] - GridGlobalStart[ai]);
        R.hi_i = (int)(Fx->RightFluxEndGlobalIndex[dim][ai] - GridGlobalStart[ai]);
        R.lo_j = (int)(Fx->RightFluxStartGlobalIndex[dim][aj] - GridGlobalStart[aj]);
        R.hi_j = (int)(Fx->RightFluxEndGlobalIndex[dim][aj] - GridGlobalStart[aj]);
        R.nface = (long long)(R.hi_i - R.lo_i + 1) * (R.hi_j - R.lo_j + 1);
        recs.push_back(R);
        rsub.push_back(n);
        rside.push_back(side);
      }
    }
    std::vector<int> idx(recs.size());
    for (size_t k = 0; k < idx.size(); k++) idx[k] = (int)k;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return recs[a].pos < recs[b].pos; });
    std::vector<FaceRec> sorted;
    std::vector<int> ssub, sside;
    for (int k : idx) {
      FaceRec R = recs[k];
      R.base = total;
      total += R.nface * nslot;
      sorted.push_back(R);
      ssub.push_back(rsub[k]);
      sside.push_back(rside[k]);
    }
    std::vector<int> head(N + 2, 0);
    for (int pos = 0, r = 0; pos <= N + 1; pos++) {
      while (r < (int)sorted.size() && sorted[r].pos < pos) r++;
      head[pos] = r;
    }
    std::vector<double> fout((size_t)std::max<long long>(total, 1), 0.0);
    SweepParams S;
    memset(&S, 0, sizeof(S));
    S.dim = dim; S.N = N; S.i1 = i1; S.i2 = i2; S.rank = rank;
    for (int d = 0; d < 3; d++) S.GD[d] = GD[d];
    S.cells = cells;
    S.nseg = (int)so[dim].size();
    S.seg_o = so[dim].data();
    S.seg_w = sw[dim].data();
    S.A = A.data();
    S.B = B.data();
    S.ncol = ncol;
    S.pres = pres.data();
    S.gr = GravityOn ? &accel[(size_t)dim * cells] : nullptr;
    S.flat = flat.data();
    S.tab = tabs[dim].data(); S.tab_row = tabs[rax].data(); S.tab_out = tabs[oax].data();
    S.qc = qc.data();
    S.uni[0] = uni[dim]; S.uni[1] = uni[rax]; S.uni[2] = uni[oax];
    for (int t = 0; t < NTAB; t++) { S.kt[t] = kt[dim][t]; S.kt_row[t] = kt[rax][t]; S.kt_out[t] = kt[oax][t]; }
    set_dx_inverse(S);
    S.gamma = gl->Gamma; S.eta1 = gl->DualEnergyFormalismEta1; S.eta2 = gl->DualEnergyFormalismEta2;
    S.pmin = gl->tiny_number; S.dfloor = gl->SmallRho; S.dt = g->dtFixed; S.dtp = nullptr;
    S.isteep = (int)gl->PPMSteepeningParameter; S.iflatten = (int)gl->PPMFlatteningParameter;
    S.idiff = (int)gl->PPMDiffusionParameter; S.idual = (int)gl->DualEnergyFormalism;
    S.ipresfree = (int)gl->PressureFree; S.gravity = (int)GravityOn;
    S.ifallback = (int)gl->RiemannSolverFallback;
    S.nact_out = (int)nact[oax];
    S.face_head = total > 0 ? head.data() : nullptr;
    S.faces = sorted.data();
    S.face_out = fout.data();
    S.counters = counters;
    S.fb_ok = pencil_fits<kFbTS>(S) ? 1 : 0;
    S.tp = kTP;
    S.fb_tag = 0;
    auto run_blocks = [&](long long nblk) {
      // main launch: flag pencil groups with a failing cell or an error
      std::vector<int> flagged;
      std::vector<char> seen((size_t)nblk / S.nseg + 1, 0);
      for (long long b = 0; b < nblk; b++) {
        std::fill(sm.begin(), sm.end(), NAN);
        const int t = xcd_tile((int)b, (int)nblk);  // the device's tile order
        TileCounts tc = S.uni[0] ? sweep_tile<kTP, kTS, 1>(S, t, sm.data(), *ex)
                                 : sweep_tile<kTP, kTS, 0>(S, t, sm.data(), *ex);
        counters[1] += tc.chains;
        if ((tc.bad || tc.err) && !seen[t / S.nseg]) {
          seen[t / S.nseg] = 1;
          flagged.push_back(t / S.nseg);
        }
      }
      // fallback launch: every pencil of a flagged group, whole
      for (int grp : flagged) {
        const int tb = grp % S.nblk_t1, t2 = S.t2_0 + grp / S.nblk_t1;
        for (int p = 0; p < S.tp && tb * S.tp + p < S.n_t1; p++) {
          if (!S.fb_ok) { counters[3]++; break; }
          std::fill(smfb.begin(), smfb.end(), NAN);
          TileCounts tc = S.uni[0] ? sweep_pencil_fb<kFbTS, 1>(S, tb * S.tp + p, t2, smfb.data(), *exfb)
                                   : sweep_pencil_fb<kFbTS, 0>(S, tb * S.tp + p, t2, smfb.data(), *exfb);
          counters[0] += tc.nfb;
          counters[2] += tc.err;
        }
      }
    };
    if (S.idiff == 0) {
      if (dim == 0) { S.ax_t1 = 1; S.ax_t2 = 2; }
      else if (dim == 1) { S.ax_t1 = 0; S.ax_t2 = 2; }
      else { S.ax_t1 = 0; S.ax_t2 = 1; }
      S.n_t1 = GD[S.ax_t1]; S.n_t2 = GD[S.ax_t2]; S.t2_0 = 0;
      set_strides(S);
      S.nblk_t1 = (S.n_t1 + kTP - 1) / kTP;
      run_blocks((long long)S.nseg * S.nblk_t1 * S.n_t2);
    } else {
      S.ax_t1 = rax; S.ax_t2 = oax; S.n_t1 = GD[rax]; S.n_t2 = 1;
      set_strides(S);
      S.nblk_t1 = (S.n_t1 + kTP - 1) / kTP;
      for (int k = 0; k < GD[oax]; k++) {
        S.t2_0 = k;
        run_blocks((long long)S.nseg * S.nblk_t1);
      }
    }
    std::swap(A, B);
    // scatter captured face fluxes (hipPPMSaveSubgridFluxes rules)
    for (size_t r = 0; r < sorted.size(); r++) {
      const FaceRec &R = sorted[r];
      enzo_fluxes *Fx = SubgridFluxes[ssub[r]];
      for (int s = 0; s < nslot; s++) {
        if (fnum[s] < 0) continue;
        if (s >= S_VX && s <= S_VZ && (s - S_VX) != dim && nact[s - S_VX] <= 1) continue;
        double *out = sside[r] == 0 ? Fx->LeftFluxes[fnum[s]][dim] : Fx->RightFluxes[fnum[s]][dim];
        if (out) memcpy(out, &fout[R.base + (size_t)s * R.nface], sizeof(double) * R.nface);
      }
    }
  }
  for (int s = 0; s < nslot; s++)
    if (fnum[s] >= 0) memcpy(g->BaryonField[fnum[s]], &A[(size_t)s * cells], sizeof(double) * cells);
  delete ex;
  delete exfb;
  for (int k = 0; k < 4; k++) counters_out[k] = (int64_t)counters[k];
  if (counters[3]) return ENZO_PPM_ERR_NOT_SUPPORTED;
  if (counters[2]) return ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}
