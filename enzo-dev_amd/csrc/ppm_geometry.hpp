// ppm_geometry.hpp -- host-side geometry of the sweep: per-axis coefficient
// tables and the split of each pencil into tile segments.
#pragma once

#include <algorithm>
#include <vector>

#include "ppm_sweep.hpp"

namespace ppm {

// Coefficient tables of one axis from its cell widths, with the reference's
// expressions (inteuler.F:140-158 grid coefficients, :166-175 steepening
// sums, calcdiss.F:382-392 diffusion factors, dx2i = 0.5/dx).
// t holds NTAB * N doubles.
inline void build_tables(const double *dx, int N, double *t) {
  for (size_t k = 0; k < (size_t)NTAB * N; k++) t[k] = 0.0;
  auto X = [&](int i) { return dx[i - 1]; };
  auto S = [&](int tt, int i) -> double & { return t[(size_t)tt * N + (i - 1)]; };
  for (int i = 1; i <= N; i++) {
    S(T_DX, i) = X(i);
    if (i - 1 >= 1 && i + 1 <= N) {
      double qa = X(i) / (X(i - 1) + X(i) + X(i + 1));
      S(T_C1, i) = qa * (2.0 * X(i - 1) + X(i)) / (X(i + 1) + X(i));
      S(T_C2, i) = qa * (2.0 * X(i + 1) + X(i)) / (X(i - 1) + X(i));
      S(T_SDXP, i) = X(i + 1) + X(i);
      S(T_SDXM, i) = X(i) + X(i - 1);
      S(T_QA3, i) = X(i - 1) + X(i) + X(i + 1);
      S(T_DDEN, i) = 0.5 * (X(i + 1) + X(i - 1)) + X(i);
      double bm = 0.5 * (X(i - 1) + X(i)), b0 = 0.5 * (X(i) + X(i + 1));
      S(T_STNUM, i) = bm * (bm * bm) + b0 * (b0 * b0);
      S(T_STDEN, i) = b0 + bm;
    }
    if (i - 2 >= 1 && i + 1 <= N) {
      double qa = X(i - 2) + X(i - 1) + X(i) + X(i + 1);
      double qb = X(i - 1) / (X(i - 1) + X(i));
      double qc = (X(i - 2) + X(i - 1)) / (2.0 * X(i - 1) + X(i));
      double qd = (X(i + 1) + X(i)) / (2.0 * X(i) + X(i - 1));
      qb = qb + 2.0 * X(i) * qb / qa * (qc - qd);
      S(T_C3, i) = 1.0 - qb;
      S(T_C4, i) = qb;
      S(T_C5, i) = X(i) / qa * qd;
      S(T_C6, i) = -X(i - 1) / qa * qc;
    }
    S(T_DX2I, i) = 0.5 / X(i);
    if (i - 1 >= 1) S(T_DIFFX, i) = 0.25 * (X(i) + X(i - 1));
  }
}
inline void build_tables(const double *dx, int N, std::vector<double> &t) {
  t.assign((size_t)NTAB * N, 0.0);
  build_tables(dx, N, t.data());
}

// If every width of the axis is the same double, every defined table entry is
// the same double too: returns true and the entries (taken at an interior
// cell) in k[NTAB].
inline bool uniform_axis(const double *dx, int N, const double *t, double *k) {
  for (int i = 1; i < N; i++)
    if (dx[i] != dx[0]) return false;
  int c = N >= 5 ? N / 2 + 1 : 0;  // 1-based interior cell with every entry defined
  for (int tt = 0; tt < NTAB; tt++) k[tt] = c ? t[(size_t)tt * N + (c - 1)] : 0.0;
  if (!c) k[T_DX] = dx[0];
  return true;
}
inline bool uniform_axis(const double *dx, int N, const std::vector<double> &t, double *k) {
  return uniform_axis(dx, N, t.data(), k);
}

// Split cells 1..N into segments owned by one tile each; every tile must
// fit its interfaces in TS threads and its loaded cells in LW.
// Raw form: o[] and w[] hold up to cap entries; returns the number of
// segments, 0 when none fits (or more than cap would be needed).
template <int TP, int TS>
inline int build_segments(int N, int i1, int i2, int *o, int *w, int cap) {
  for (int nseg = (N + TS - 8) / (TS - 7); nseg <= N && nseg <= cap; nseg++) {
    int base = N / nseg, rem = N % nseg, start = 1;
    bool ok = true;
    for (int s = 0; s < nseg; s++) {
      int ww = base + (s < rem ? 1 : 0);
      if (ww < 1) { ok = false; break; }
      int IA = std::max(start - 1, i1), IB = std::min(start + ww + 1, i2 + 1);
      int NI = IB - IA + 1;
      int L0 = std::min(IA - 3, std::max(start - 1, 1));
      int L1 = std::max(IB + 2, std::min(start + ww, N));
      // one slope item per thread (cells [IA-2, IB+1]) and the loaded cells in LW
      if (NI + 4 > TS || L1 - L0 + 1 > Cfg<TP, TS>::LW || L0 < 1 || L1 > N) { ok = false; break; }
      o[s] = start;
      w[s] = ww;
      start += ww;
    }
    if (ok) return nseg;
  }
  return 0;
}
template <int TP, int TS>
inline bool build_segments(int N, int i1, int i2, std::vector<int> &o, std::vector<int> &w) {
  o.assign(N, 0);
  w.assign(N, 0);
  const int n = build_segments<TP, TS>(N, i1, i2, o.data(), w.data(), N);
  o.resize(n);
  w.resize(n);
  return n > 0;
}

}  // namespace ppm
