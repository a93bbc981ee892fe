/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement, in plain C (FP64 reals, int64 integers: the reference's
 * default precision-64 / integers-64 build), of the per-slice Fortran kernels
 * of Enzo's PPM "direct Eulerian" hydro step.  Every routine keeps the
 * reference's Fortran calling convention (all arguments by reference,
 * column-major (idim,jdim[,ncolor]) slices, 1-based inclusive index bounds)
 * so that the same slice driver (ppm_driver_oracle.c) can run either these
 * functions or the flang-built reference objects, and so that each routine
 * can be compared one-to-one with its reference counterpart.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py may load
 * this code, and only as the checker.  The product (the HIP library) never
 * links or calls it.
 *
 * Arithmetic follows the reference expression by expression (operand order,
 * parenthesisation, no contraction: build with -ffp-contract=off) because
 * the PPM limiters, steepening and dual-energy switches branch on the
 * results.  Fortran intrinsics are restated with the semantics that flang
 * lowers them to:  MAX(a,b) = a>b ? a : b,  MIN(a,b) = a<b ? a : b
 * (multi-argument forms fold left to right), SIGN(x,y) = copysign(|x|,y),
 * x**2 = x*x, x**3 = x*(x*x).
 *
 * Supported options (the ones the reference path uses):
 *   calcdiss   iflatten = 0,1,2,3; idiff = 0,1.   idiff = 2 reads sigma/kappa
 *              outside the range calcdiss.F fills (calcdiss.F:602-607 with
 *              :520-528) and is rejected.
 *   inteuler   iconsrec = 0, iposrec = 0 (conservative / positive
 *              reconstruction are off on this path and rejected).
 *   flux_twoshock with the per-cell flux_hll fallback.
 *
 * Parity status: every routine here is pinned bit-for-bit against the
 * reference Fortran built with flang (oracle/Makefile, target ref) on
 * seeded slices and whole grids, against committed golden fixtures made from
 * that build, and against closed forms (see tests/).  The only exception is
 * the value the reference leaves undefined: flux_hll called on one cell
 * (flux_twoshock.F:444-450) evaluates ges(i+1) from uninitialised locals
 * sl(i+2), gesl(i+2) (flux_hll.F:302-307).  Here those two locals are given
 * the values flux_hll would compute at interface i+2, which makes the result
 * defined; cells that take the fallback are counted and reported separately.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t I;   /* INTG_PREC under LARGE_INTS      (fortran_types.def:43-49) */
typedef double R;    /* R_PREC under CONFIG_BFLOAT_8    (fortran_types.def:28-33) */

#define TINY 1e-20          /* tiny  (fortran_types.def:7-12, PFLOAT_8) */
#define COLOR_FLOOR 1e-35   /* color_floor (fortran.def:8) */
#define IJKN 1031           /* MAX_ANY_SINGLE_DIRECTION (fortran.def:1) */
#define MAXCOLOR 45         /* MAX_COLOR (fortran.def:2) */

/* Error/fallback bookkeeping shared with the driver (per thread). */
_Thread_local I orc_error_code = 0;      /* 0 ok; nonzero: the reference would abort */
_Thread_local I orc_fallback_count = 0;  /* cells recomputed with flux_hll */

static inline R fmx(R a, R b) { return a > b ? a : b; }
static inline R fmn(R a, R b) { return a < b ? a : b; }
static inline R sgn1(R b) { return copysign(1.0, b); }
static inline R sgn2(R b) { return copysign(2.0, b); }

/* 1-based column-major accessors */
#define S2(a, i, j) (a)[((i) - 1) + ((j) - 1) * idim]
#define S3(a, i, j, n) (a)[((i) - 1) + ((j) - 1) * idim + ((n) - 1) * idim * jdim]
#define V3(a, i, j, k) (a)[((i) - 1) + ((j) - 1) * dimx + ((k) - 1) * dimx * dimy]
#define F1(a, i) (a)[(i) - 1]

/* ------------------------------------------------------------------ */
/* calcdiss.F:17-623 -- flattening (CW84 A1-A2, A4-A6, A7-A10) and the
 * idiff=1 diffusion coefficient K max(-div u, 0) (CW84 A3).            */
void orc_calcdiss(const R *dslice, const R *eslice, const R *uslice,
                  const R *v, const R *w, const R *pslice,
                  const R *dx, const R *dy, const R *dz,
                  const I *pidim, const I *pjdim, const I *pkdim,
                  const I *pi1, const I *pi2, const I *pj1, const I *pj2,
                  const I *pk, const I *pnzz, const I *pidir,
                  const I *pdimx, const I *pdimy, const I *pdimz,
                  const R *pdt, const R *pgamma, const I *pidiff,
                  const I *piflatten, R *diffcoef, R *flatten) {
  const I idim = *pidim, jdim = *pjdim, kdim = *pkdim, i1 = *pi1, i2 = *pi2;
  const I j1 = *pj1, j2 = *pj2, k = *pk, nzz = *pnzz, idir = *pidir;
  const I dimx = *pdimx, dimy = *pdimy, dimz = *pdimz, idiff = *pidiff;
  const I iflatten = *piflatten;
  const R gamma = *pgamma;
  (void)pdt;
  /* calcdiss.F:143-151 */
  const R epsilon = 0.33, kappa1 = 2.0, kappa2 = 0.01, Kparam = 0.1;
  const R omega1 = 0.75, omega2 = 10.0, sigma1 = 0.5, sigma2 = 1.0;
  R wflag[IJKN + 8], vdiff1[IJKN + 8], wdiff1[IJKN + 8], flattemp[IJKN + 8];
  R di[IJKN + 8], omega[IJKN + 8], kappa[IJKN + 8], sigma[IJKN + 8];
  I i, j;

  if (idiff == 2) { orc_error_code = 2; return; }

  const I nyz = j2 - j1 + 1; /* :163 */
  const int lk1 = (1 < k && k < kdim);
  const int lkx1 = (1 < k && k < dimx), lky1 = (1 < k && k < dimy);
  const int lkz1 = (1 < k && k < dimz);
  (void)lk1; (void)lky1;

  for (j = j1; j <= j2; j++) {
    const int lj1 = (1 < j && j < jdim);
    const int ljx1 = (1 < j && j < dimx), ljy1 = (1 < j && j < dimy);
    const int ljz1 = (1 < j && j < dimz);

    /* CW84 eq. A1 shock flag (:193-202) */
    for (i = i1 - 2; i <= i2 + 2; i++) {
      R qb = fabs(S2(pslice, i + 1, j) - S2(pslice, i - 1, j)) /
             fmn(S2(pslice, i + 1, j), S2(pslice, i - 1, j));
      if (qb > epsilon && S2(uslice, i - 1, j) > S2(uslice, i + 1, j))
        wflag[i] = 1.0;
      else
        wflag[i] = 0.0;
    }

    /* transverse velocity differences for idiff=1 (:207-288), read from the
       full 3-D v and w arrays as the reference does */
    if (idiff == 1) {
      if (idir == 1) {
        if (dimy > 1 && ljy1)
          for (i = i1 - 2; i <= i2 + 2; i++)
            vdiff1[i] = (V3(v, i, j - 1, k) + V3(v, i - 1, j - 1, k)) -
                        (V3(v, i, j + 1, k) + V3(v, i - 1, j + 1, k));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) vdiff1[i] = 0.0;
        if (dimz > 1 && lkz1)
          for (i = i1 - 2; i <= i2 + 2; i++)
            wdiff1[i] = (V3(w, i, j, k - 1) + V3(w, i - 1, j, k - 1)) -
                        (V3(w, i, j, k + 1) + V3(w, i - 1, j, k + 1));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) wdiff1[i] = 0.0;
      } else if (idir == 2) {
        if (dimz > 1 && ljz1)
          for (i = i1 - 2; i <= i2 + 2; i++)
            vdiff1[i] = (V3(v, k, i, j - 1) + V3(v, k, i - 1, j - 1)) -
                        (V3(v, k, i, j + 1) + V3(v, k, i - 1, j + 1));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) vdiff1[i] = 0.0;
        if (dimx > 1 && lkx1)
          for (i = i1 - 2; i <= i2 + 2; i++)
            wdiff1[i] = (V3(w, k - 1, i, j) + V3(w, k - 1, i - 1, j)) -
                        (V3(w, k + 1, i, j) + V3(w, k + 1, i - 1, j));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) wdiff1[i] = 0.0;
      } else if (idir == 3) {
        if (dimx > 1 && ljx1)
          for (i = i1 - 2; i <= i2 + 2; i++)
            vdiff1[i] = (V3(v, j - 1, k, i) + V3(v, j - 1, k, i - 1)) -
                        (V3(v, j + 1, k, i) + V3(v, j + 1, k, i - 1));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) vdiff1[i] = 0.0;
        if (dimy > 1 && (1 < k && k < dimy))
          for (i = i1 - 2; i <= i2 + 2; i++)
            wdiff1[i] = (V3(w, j, k - 1, i) + V3(w, j, k - 1, i - 1)) -
                        (V3(w, j, k + 1, i) + V3(w, j, k + 1, i - 1));
        else
          for (i = i1 - 2; i <= i2 + 2; i++) wdiff1[i] = 0.0;
      }
    }

    /* (B) diffusion coefficient (:377-404) */
    if (idiff == 1) {
      for (i = i1; i <= i2 + 1; i++)
        S2(diffcoef, i, j) = S2(uslice, i - 1, j) - S2(uslice, i, j);
      if (nyz > 1 && lj1)
        for (i = i1; i <= i2 + 1; i++)
          S2(diffcoef, i, j) =
              S2(diffcoef, i, j) +
              (0.25 * (F1(dx, i) + F1(dx, i - 1)) /
               (0.5 * (F1(dy, j + 1) + F1(dy, j - 1)) + F1(dy, j))) *
                  vdiff1[i];
      if (nzz > 1 && lk1)
        for (i = i1; i <= i2 + 1; i++)
          S2(diffcoef, i, j) =
              S2(diffcoef, i, j) +
              (0.25 * (F1(dx, i) + F1(dx, i - 1)) /
               (0.5 * (F1(dz, k + 1) + F1(dz, k - 1)) + F1(dz, k))) *
                  wdiff1[i];
      for (i = i1; i <= i2 + 1; i++)
        S2(diffcoef, i, j) = Kparam * fmx(0.0, S2(diffcoef, i, j));
    }

    /* (E) flattening, eqs A1-A2 (:409-442) */
    if (iflatten == 1) {
      for (i = i1 - 1; i <= i2 + 1; i++) {
        R qa;
        if (fabs(S2(pslice, i + 2, j) - S2(pslice, i - 2, j)) /
                fmn(S2(pslice, i + 2, j), S2(pslice, i - 2, j)) <
            epsilon)
          qa = 1.0;
        else
          qa = (S2(pslice, i + 1, j) - S2(pslice, i - 1, j)) /
               (S2(pslice, i + 2, j) - S2(pslice, i - 2, j));
        flattemp[i] = fmn(1.0, (qa - omega1) * omega2 * wflag[i]);
        flattemp[i] = fmx(0.0, flattemp[i]);
      }
      flattemp[i1 - 2] = flattemp[i1 - 1];
      flattemp[i2 + 2] = flattemp[i2 + 1];
      for (i = i1 - 1; i <= i2 + 1; i++) {
        if (S2(pslice, i + 1, j) - S2(pslice, i - 1, j) < 0.0)
          S2(flatten, i, j) = fmx(flattemp[i], flattemp[i + 1]);
        else
          S2(flatten, i, j) = fmx(flattemp[i], flattemp[i - 1]);
      }
    }

    /* (F) flattening for Lagrangean hydro, eqs A4-A6 (:447-485) */
    if (iflatten == 2) {
      for (i = i1 - 3; i <= i2 + 3; i++) di[i] = 1.0 / S2(dslice, i, j);
      for (i = i1 - 1; i <= i2 + 1; i++) {
        I is = (I)((R)i + sgn2(S2(pslice, i + 1, j) - S2(pslice, i - 1, j)));
        omega[i] = fmx(0.0, omega1 * (omega2 - (S2(pslice, i + 1, j) - S2(pslice, i - 1, j)) /
                                                   (S2(pslice, i + 2, j) - S2(pslice, i - 2, j))));
        R Z = sqrt((fmx(S2(pslice, i + 2, j), S2(pslice, i - 2, j)) +
                    0.5 * (S2(pslice, i + 2, j) + S2(pslice, i - 2, j)) * (gamma - 1.0)) /
                   fmx(di[i + 2], di[i - 2]));
        R kappa_tilde = (Z + sqrt(gamma * S2(pslice, is, j) * S2(dslice, is, j))) / Z;
        kappa[i] = fmx(0.0, (kappa_tilde - kappa1) / (kappa_tilde + kappa2));
        flattemp[i] = fmn(wflag[i] * omega[i], kappa[i]);
      }
      flattemp[i1 - 2] = flattemp[i1 - 1];
      flattemp[i2 + 2] = flattemp[i2 + 1];
      for (i = i1 - 1; i <= i2 + 1; i++)
        S2(flatten, i, j) = fmx(fmx(flattemp[i - 1], flattemp[i]), flattemp[i + 1]);
    }

    /* (G) multidimensional flattening, eqs A7-A9 (:489-566) */
    if (iflatten == 3) {
      for (i = i1 - 3; i <= i2 + 3; i++) di[i] = 1.0 / S2(dslice, i, j);
      for (i = i1 - 1; i <= i2 + 1; i++) {
        R dp1 = S2(pslice, i + 1, j) - S2(pslice, i - 1, j);
        R dp2 = S2(pslice, i + 2, j) - S2(pslice, i - 2, j);
        R de1 = S2(eslice, i + 1, j) - S2(eslice, i - 1, j);
        R de2 = S2(eslice, i + 2, j) - S2(eslice, i - 2, j);
        R dpp = 0.0, dee = 0.0;
        if (dp2 != 0.0) dpp = dp1 / dp2;
        if (de2 != 0.0) dee = de1 / de2;
        R omega_tilde = fmx(dpp, dee);
        I ism = (I)((R)i + sgn2(dp1));
        I isp = (I)((R)i - sgn2(dp1));
        R s = -sgn1(dp1);
        if (dp1 == 0.0) s = 0.0;
        R sigma_tilde = wflag[i] * fabs(dp2) / fmn(S2(pslice, i + 2, j), S2(pslice, i - 2, j));
        sigma[i] = fmx(0.0, (sigma_tilde - sigma1) / (sigma_tilde + sigma2));
        omega[i] = fmx(0.0, omega2 * (omega_tilde - omega1));
        R Z = sqrt((fmx(S2(pslice, i + 2, j), S2(pslice, i - 2, j)) +
                    0.5 * (S2(pslice, i + 2, j) + S2(pslice, i - 2, j)) * (gamma - 1.0)) /
                   fmx(di[i + 2], di[i - 2]));
        R ZE = s * Z / S2(dslice, ism, j) + S2(uslice, ism, j) + TINY;
        R cj2s = sqrt(gamma * S2(pslice, isp, j) / S2(dslice, isp, j));
        R kappa_tilde = fabs((ZE - S2(uslice, isp, j) + s * cj2s) / ZE);
        kappa[i] = fmx(0.0, (kappa_tilde - kappa1) / (kappa_tilde + kappa2));
        flattemp[i] = fmn(fmn(kappa[i], wflag[i] * omega[i]), wflag[i] * sigma[i]);
      }
    }

    /* A.2 maximum for iflatten = 2, 3 (:568-585) */
    if (iflatten == 2 || iflatten == 3) {
      flattemp[i1 - 2] = flattemp[i1 - 1];
      flattemp[i2 + 2] = flattemp[i2 + 1];
      for (i = i1 - 1; i <= i2 + 1; i++) {
        if (S2(pslice, i + 1, j) - S2(pslice, i - 1, j) < 0.0)
          S2(flatten, i, j) = fmx(flattemp[i], flattemp[i + 1]);
        else
          S2(flatten, i, j) = fmx(flattemp[i], flattemp[i - 1]);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* intvar.F:5-178 -- PPM interpolation of one variable along one line.   */
static void orc_intvar(const R *q, I i1, I i2, I isteep, const R *steepen,
                       I iflatten, const R *flat, const R *c1, const R *c2,
                       const R *c3, const R *c4, const R *c5, const R *c6,
                       const R *char1, const R *char2, const R *c0, R *dq,
                       R *ql, R *qr, R *q6, R *qla, R *qra, R *ql0, R *qr0,
                       I idim) {
  /* arrays here are 1-based: q[i], flat[i] already offset by the caller */
  const R ft = 4.0 / 3.0;
  I i;
  for (i = 1; i <= idim; i++) { dq[i] = 0.0; ql[i] = 0.0; qr[i] = 0.0; q6[i] = 0.0; } /* :69-72 */

  /* average slopes, centred / van Leer, monotonised (:82-94) */
  for (i = i1 - 2; i <= i2 + 2; i++) {
    R qplus = q[i + 1] - q[i];
    R qmnus = q[i] - q[i - 1];
    if (qplus * qmnus > 0) {
      R qcent = c1[i] * qplus + c2[i] * qmnus;
      R qvanl = 2.0 * qplus * qmnus / (qmnus + qplus);
      R temp1 = fmn(fmn(fmn(fabs(qcent), fabs(qvanl)), 2.0 * fabs(qmnus)), 2.0 * fabs(qplus));
      dq[i] = temp1 * sgn1(qcent);
    } else {
      dq[i] = 0.0;
    }
  }
  /* interface values (:98-102) */
  for (i = i1 - 1; i <= i2 + 2; i++) {
    ql[i] = c3[i] * q[i - 1] + c4[i] * q[i] + c5[i] * dq[i - 1] + c6[i] * dq[i];
    qr[i - 1] = ql[i];
  }
  /* steepening (:106-113) */
  if (isteep != 0) {
    for (i = i1 - 1; i <= i2 + 1; i++) {
      ql[i] = (1.0 - steepen[i]) * ql[i] + steepen[i] * (q[i - 1] + 0.5 * dq[i - 1]);
      qr[i] = (1.0 - steepen[i]) * qr[i] + steepen[i] * (q[i + 1] - 0.5 * dq[i + 1]);
    }
  }
  /* monotonise (:117-131) */
  for (i = i1 - 1; i <= i2 + 1; i++) {
    R temp1 = (qr[i] - q[i]) * (q[i] - ql[i]);
    R temp2 = qr[i] - ql[i];
    R temp3 = 6.0 * (q[i] - 0.5 * (qr[i] + ql[i]));
    if (temp1 <= 0.0) { ql[i] = q[i]; qr[i] = q[i]; }
    R temp22 = temp2 * temp2;
    R temp23 = temp2 * temp3;
    if (temp22 < temp23) ql[i] = 3.0 * q[i] - 2.0 * qr[i];
    if (temp22 < -temp23) qr[i] = 3.0 * q[i] - 2.0 * ql[i];
  }
  /* flattening (:135-140) */
  if (iflatten != 0) {
    for (i = i1 - 1; i <= i2 + 1; i++) {
      ql[i] = q[i] * flat[i] + ql[i] * (1.0 - flat[i]);
      qr[i] = q[i] * flat[i] + qr[i] * (1.0 - flat[i]);
    }
  }
  /* CHECK_LR (:145-153) */
  for (i = i1 - 1; i <= i2 + 2; i++) {
    ql[i] = fmx(fmn(q[i], q[i - 1]), ql[i]);
    ql[i] = fmn(fmx(q[i], q[i - 1]), ql[i]);
    qr[i] = fmx(fmn(q[i], q[i + 1]), qr[i]);
    qr[i] = fmn(fmx(q[i], q[i + 1]), qr[i]);
  }
  /* parabola (:157-160) */
  for (i = i1 - 1; i <= i2 + 1; i++) {
    q6[i] = 6.0 * (q[i] - 0.5 * (ql[i] + qr[i]));
    dq[i] = qr[i] - ql[i];
  }
  /* domain-of-dependence averages (:162-174) */
  for (i = i1; i <= i2 + 1; i++) {
    qla[i] = qr[i - 1] - char1[i - 1] * (dq[i - 1] - (1.0 - ft * char1[i - 1]) * q6[i - 1]);
    qra[i] = ql[i] + char2[i] * (dq[i] + (1.0 - ft * char2[i]) * q6[i]);
  }
  for (i = i1; i <= i2 + 1; i++) {
    ql0[i] = qr[i - 1] - c0[i - 1] * (dq[i - 1] - (1.0 - ft * c0[i - 1]) * q6[i - 1]);
    qr0[i] = ql[i] - c0[i] * (dq[i] + (1.0 + ft * c0[i]) * q6[i]);
  }
}

/* ------------------------------------------------------------------ */
/* inteuler.F:5-566 -- left/right Eulerian interface states.            */
void orc_inteuler(const R *dslice, const R *pslice, const I *pgravity,
                  const R *grslice, const R *geslice, const R *uslice,
                  const R *vslice, const R *wslice, const R *dxi,
                  const R *flatten, const I *pidim, const I *pjdim,
                  const I *pi1, const I *pi2, const I *pj1, const I *pj2,
                  const I *pidual, const R *peta1, const R *peta2,
                  const I *pisteep, const I *piflatten, const I *piconsrec,
                  const I *piposrec, const R *pdt, const R *pgamma,
                  const I *pipresfree, R *dls, R *drs, R *pls, R *prs,
                  R *gels, R *gers, R *uls, R *urs, R *vls, R *vrs, R *wls,
                  R *wrs, const I *pncolor, const R *colslice, R *colls,
                  R *colrs) {
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const I gravity = *pgravity, idual = *pidual, isteep = *pisteep;
  const I iflatten = *piflatten, ipresfree = *pipresfree, ncolor = *pncolor;
  const R eta2 = *peta2, dt = *pdt, gamma = *pgamma;
  const R ft = 4.0 / 3.0;
  (void)peta1;
  I i, j, ic;
  if (*piconsrec != 0 || *piposrec != 0) { orc_error_code = 3; return; }
  if (ncolor > MAXCOLOR) { orc_error_code = 4; return; }

  /* 1-based local work arrays (ijkn) */
#define LA(name) R name[IJKN + 2]
  LA(c1); LA(c2); LA(c3); LA(c4); LA(c5); LA(c6); LA(dx2i); LA(steepen);
  LA(d2d); LA(dxb); LA(cs); LA(char1); LA(char2); LA(cm); LA(c0); LA(cp);
  LA(tmp1); LA(tmp2); LA(tmp3); LA(tmp4); LA(dla); LA(dra); LA(dl0); LA(dr0);
  LA(dp); LA(pl); LA(pr); LA(p6); LA(pla); LA(pra); LA(pl0); LA(pr0);
  LA(du); LA(ul); LA(ur); LA(u6); LA(ula); LA(ura); LA(ul0); LA(ur0);
  LA(dv); LA(vl); LA(vr); LA(v6); LA(vla); LA(vra); LA(vl0); LA(vr0);
  LA(dw); LA(wl); LA(wr); LA(w6); LA(wla); LA(wra); LA(wl0); LA(wr0);
  LA(dge); LA(gel); LA(ger); LA(ge6); LA(gela); LA(gera); LA(gel0); LA(ger0);
  LA(plm); LA(prm); LA(plp); LA(prp); LA(ulm); LA(urm); LA(ulp); LA(urp);
  LA(cla); LA(cra); LA(betalp); LA(betalm); LA(betal0); LA(betarp);
  LA(betarm); LA(betar0);
  LA(cdq); LA(cql); LA(cqr); LA(cq6);
#undef LA
  /* colour work: colla/colra/coll0/colr0 per colour */
  R *colw = (R *)malloc(sizeof(R) * (size_t)(IJKN + 2) * 4 * (size_t)(ncolor > 0 ? ncolor : 1));
  /* the Fortran locals gela..ger0 are undefined when idual = 0 but are still
     copied into gels/gers (:424-439); keep them defined (zero) here */
  for (i = 0; i < IJKN + 2; i++) { gela[i] = gera[i] = gel0[i] = ger0[i] = 0.0; }

  const R *dxv = dxi - 1; /* 1-based */
  /* grid coefficients (:140-158) */
  for (i = i1 - 2; i <= i2 + 2; i++) {
    R qa = dxv[i] / (dxv[i - 1] + dxv[i] + dxv[i + 1]);
    c1[i] = qa * (2.0 * dxv[i - 1] + dxv[i]) / (dxv[i + 1] + dxv[i]);
    c2[i] = qa * (2.0 * dxv[i + 1] + dxv[i]) / (dxv[i - 1] + dxv[i]);
  }
  for (i = i1 - 1; i <= i2 + 2; i++) {
    R qa = dxv[i - 2] + dxv[i - 1] + dxv[i] + dxv[i + 1];
    R qb = dxv[i - 1] / (dxv[i - 1] + dxv[i]);
    R qc = (dxv[i - 2] + dxv[i - 1]) / (2.0 * dxv[i - 1] + dxv[i]);
    R qd = (dxv[i + 1] + dxv[i]) / (2.0 * dxv[i] + dxv[i - 1]);
    qb = qb + 2.0 * dxv[i] * qb / qa * (qc - qd);
    c3[i] = 1.0 - qb;
    c4[i] = qb;
    c5[i] = dxv[i] / qa * qd;
    c6[i] = -dxv[i - 1] / qa * qc;
    dx2i[i] = 0.5 / dxv[i];
  }

  for (j = j1; j <= j2; j++) {
    const R *d = &S2(dslice, 1, j) - 1; /* 1-based pencil views */
    const R *p = &S2(pslice, 1, j) - 1;
    const R *u = &S2(uslice, 1, j) - 1;
    const R *vv = &S2(vslice, 1, j) - 1;
    const R *ww = &S2(wslice, 1, j) - 1;
    const R *gr = gravity == 1 ? &S2(grslice, 1, j) - 1 : NULL;
    /* intvar receives `flatten`, the base of the (idim,jdim) array, not
       flatten(1,j) (inteuler.F:241-289): every line j of the slice is
       flattened with the coefficients of line 1. */
    const R *fl = flatten - 1;

    /* steepening coefficients (:166-193) */
    if (isteep != 0) {
      for (i = i1 - 2; i <= i2 + 2; i++) {
        R qa = dxv[i - 1] + dxv[i] + dxv[i + 1];
        d2d[i] = (d[i + 1] - d[i]) / (dxv[i + 1] + dxv[i]);
        d2d[i] = (d2d[i] - (d[i] - d[i - 1]) / (dxv[i] + dxv[i - 1])) / qa;
        dxb[i] = 0.5 * (dxv[i] + dxv[i + 1]);
      }
      for (i = i1 - 1; i <= i2 + 1; i++) {
        R qc = fabs(d[i + 1] - d[i - 1]) - 0.01 * fmn(fabs(d[i + 1]), fabs(d[i - 1]));
        R s1 = (d2d[i - 1] - d2d[i + 1]) *
               (dxb[i - 1] * (dxb[i - 1] * dxb[i - 1]) + dxb[i] * (dxb[i] * dxb[i])) /
               ((dxb[i] + dxb[i - 1]) * (d[i + 1] - d[i - 1] + TINY));
        if (d2d[i + 1] * d2d[i - 1] > 0.0) s1 = 0.0;
        if (qc <= 0.0) s1 = 0.0;
        R s2 = fmx(0.0, fmn(20.0 * (s1 - 0.05), 1.0));
        R qa = fabs(d[i + 1] - d[i - 1]) / fmn(d[i + 1], d[i - 1]);
        R qb = fabs(p[i + 1] - p[i - 1]) / fmn(p[i + 1], p[i - 1]);
        if (gamma * 0.1 * qa >= qb)
          steepen[i] = s2;
        else
          steepen[i] = 0.0;
      }
    }
    /* characteristic distances (:197-208) */
    for (i = 1; i <= idim; i++) {
      cs[i] = sqrt(gamma * p[i] / d[i]);
      if (ipresfree == 1) cs[i] = TINY;
      if (i >= i1 - 1 && i <= i2 + 2) {
        char1[i] = fmx(0.0, dt * (u[i] + cs[i])) * dx2i[i];
        char2[i] = fmx(0.0, -dt * (u[i] - cs[i])) * dx2i[i];
      }
    }
    for (i = i1 - 2; i <= i2 + 2; i++) {
      cm[i] = dt * (u[i] - cs[i]) * dx2i[i];
      c0[i] = dt * (u[i]) * dx2i[i];
      cp[i] = dt * (u[i] + cs[i]) * dx2i[i];
    }
    /* interpolation of each variable (:241-263, :281-294) */
    orc_intvar(d, i1, i2, isteep, steepen, iflatten, fl, c1, c2, c3, c4, c5, c6,
               char1, char2, c0, tmp1, tmp2, tmp3, tmp4, dla, dra, dl0, dr0, idim);
    orc_intvar(p, i1, i2, 0, steepen, iflatten, fl, c1, c2, c3, c4, c5, c6,
               char1, char2, c0, dp, pl, pr, p6, pla, pra, pl0, pr0, idim);
    orc_intvar(u, i1, i2, 0, steepen, iflatten, fl, c1, c2, c3, c4, c5, c6,
               char1, char2, c0, du, ul, ur, u6, ula, ura, ul0, ur0, idim);
    orc_intvar(vv, i1, i2, 0, steepen, iflatten, fl, c1, c2, c3, c4, c5, c6,
               char1, char2, c0, dv, vl, vr, v6, vla, vra, vl0, vr0, idim);
    orc_intvar(ww, i1, i2, 0, steepen, iflatten, fl, c1, c2, c3, c4, c5, c6,
               char1, char2, c0, dw, wl, wr, w6, wla, wra, wl0, wr0, idim);
    if (idual == 1)
      orc_intvar(&S2(geslice, 1, j) - 1, i1, i2, 0, steepen, iflatten, fl, c1, c2, c3,
                 c4, c5, c6, char1, char2, c0, dge, gel, ger, ge6, gela, gera, gel0,
                 ger0, idim);
    for (ic = 1; ic <= ncolor; ic++) {
      R *cla_ = colw + (size_t)(ic - 1) * 4 * (IJKN + 2);
      R *cra_ = cla_ + (IJKN + 2), *cl0_ = cra_ + (IJKN + 2), *cr0_ = cl0_ + (IJKN + 2);
      orc_intvar(&S3(colslice, 1, j, ic) - 1, i1, i2, 0, steepen, iflatten, fl, c1,
                 c2, c3, c4, c5, c6, char1, char2, c0, cdq, cql, cqr, cq6, cla_,
                 cra_, cl0_, cr0_, idim);
    }

    /* averages over the characteristic domain of dependence (3.5) (:304-317) */
    for (i = i1; i <= i2 + 1; i++) {
      plm[i] = pr[i - 1] - cm[i - 1] * (dp[i - 1] - (1.0 - ft * cm[i - 1]) * p6[i - 1]);
      prm[i] = pl[i] - cm[i] * (dp[i] + (1.0 + ft * cm[i]) * p6[i]);
      plp[i] = pr[i - 1] - cp[i - 1] * (dp[i - 1] - (1.0 - ft * cp[i - 1]) * p6[i - 1]);
      prp[i] = pl[i] - cp[i] * (dp[i] + (1.0 + ft * cp[i]) * p6[i]);
    }
    for (i = i1; i <= i2 + 1; i++) {
      ulm[i] = ur[i - 1] - cm[i - 1] * (du[i - 1] - (1.0 - ft * cm[i - 1]) * u6[i - 1]);
      urm[i] = ul[i] - cm[i] * (du[i] + (1.0 + ft * cm[i]) * u6[i]);
      ulp[i] = ur[i - 1] - cp[i - 1] * (du[i - 1] - (1.0 - ft * cp[i - 1]) * u6[i - 1]);
      urp[i] = ul[i] - cp[i] * (du[i] + (1.0 + ft * cp[i]) * u6[i]);
    }
    /* correction terms (3.7) (:321-383) */
    for (i = i1; i <= i2 + 1; i++) {
      cla[i] = sqrt(fmx(gamma * pla[i] * dla[i], 0.0));
      cra[i] = sqrt(fmx(gamma * pra[i] * dra[i], 0.0));
    }
    for (i = i1; i <= i2 + 1; i++) {
      R f1 = 1.0 / cla[i];
      betalp[i] = (ula[i] - ulp[i]) + (pla[i] - plp[i]) * f1;
      betalm[i] = (ula[i] - ulm[i]) - (pla[i] - plm[i]) * f1;
      betal0[i] = (pla[i] - pl0[i]) * (f1 * f1) + 1.0 / dla[i] - 1.0 / dl0[i];
    }
    if (gravity == 1)
      for (i = i1; i <= i2 + 1; i++) {
        betalp[i] = betalp[i] - 0.25 * dt * (gr[i - 1] + gr[i]);
        betalm[i] = betalm[i] - 0.25 * dt * (gr[i - 1] + gr[i]);
      }
    for (i = i1; i <= i2 + 1; i++) {
      R f1 = 0.5 / cla[i];
      betalp[i] = -betalp[i] * f1;
      betalm[i] = +betalm[i] * f1;
    }
    for (i = i1; i <= i2 + 1; i++) {
      if (cp[i - 1] <= 0.0) betalp[i] = 0.0;
      if (cm[i - 1] <= 0.0) betalm[i] = 0.0;
      if (c0[i - 1] <= 0.0) betal0[i] = 0.0;
    }
    for (i = i1; i <= i2 + 1; i++) {
      R f1 = 1.0 / cra[i];
      betarp[i] = (ura[i] - urp[i]) + (pra[i] - prp[i]) * f1;
      betarm[i] = (ura[i] - urm[i]) - (pra[i] - prm[i]) * f1;
      betar0[i] = (pra[i] - pr0[i]) * (f1 * f1) + 1.0 / dra[i] - 1.0 / dr0[i];
    }
    if (gravity == 1)
      for (i = i1; i <= i2 + 1; i++) {
        betarp[i] = betarp[i] - 0.25 * dt * (gr[i - 1] + gr[i]);
        betarm[i] = betarm[i] - 0.25 * dt * (gr[i - 1] + gr[i]);
      }
    for (i = i1; i <= i2 + 1; i++) {
      R f1 = 0.5 / cra[i];
      betarp[i] = -betarp[i] * f1;
      betarm[i] = +betarm[i] * f1;
    }
    for (i = i1; i <= i2 + 1; i++) {
      if (cp[i] >= 0) betarp[i] = 0.0;
      if (cm[i] >= 0) betarm[i] = 0.0;
      if (c0[i] >= 0) betar0[i] = 0.0;
    }
    /* corrected states (3.6) (:387-396) */
    for (i = i1; i <= i2 + 1; i++) {
      S2(pls, i, j) = pla[i] + (betalp[i] + betalm[i]) * (cla[i] * cla[i]);
      S2(prs, i, j) = pra[i] + (betarp[i] + betarm[i]) * (cra[i] * cra[i]);
      S2(uls, i, j) = ula[i] + (betalp[i] - betalm[i]) * cla[i];
      S2(urs, i, j) = ura[i] + (betarp[i] - betarm[i]) * cra[i];
      S2(dls, i, j) = 1.0 / (1.0 / dla[i] - (betal0[i] + betalp[i] + betalm[i]));
      S2(drs, i, j) = 1.0 / (1.0 / dra[i] - (betar0[i] + betarp[i] + betarm[i]));
    }
    /* upwind advected states (:400-434) */
    for (i = i1; i <= i2 + 1; i++) {
      if (u[i - 1] <= 0.0) {
        S2(vls, i, j) = vla[i]; S2(wls, i, j) = wla[i]; S2(gels, i, j) = gela[i];
      } else {
        S2(vls, i, j) = vl0[i]; S2(wls, i, j) = wl0[i]; S2(gels, i, j) = gel0[i];
      }
      if (u[i] >= 0.0) {
        S2(vrs, i, j) = vra[i]; S2(wrs, i, j) = wra[i]; S2(gers, i, j) = gera[i];
      } else {
        S2(vrs, i, j) = vr0[i]; S2(wrs, i, j) = wr0[i]; S2(gers, i, j) = ger0[i];
      }
    }
    for (ic = 1; ic <= ncolor; ic++) {
      R *cla_ = colw + (size_t)(ic - 1) * 4 * (IJKN + 2);
      R *cra_ = cla_ + (IJKN + 2), *cl0_ = cra_ + (IJKN + 2), *cr0_ = cl0_ + (IJKN + 2);
      for (i = i1; i <= i2 + 1; i++) {
        if (u[i - 1] <= 0.0)
          S3(colls, i, j, ic) = cla_[i] * S2(dls, i, j) / dla[i];
        else
          S3(colls, i, j, ic) = cl0_[i] * S2(dls, i, j) / dl0[i];
        if (u[i] >= 0.0)
          S3(colrs, i, j, ic) = cra_[i] * S2(drs, i, j) / dra[i];
        else
          S3(colrs, i, j, ic) = cr0_[i] * S2(drs, i, j) / dr0[i];
      }
    }
    /* dual-energy override of the corrections (:493-521) */
    if (idual == 1) {
      for (i = i1; i <= i2 + 1; i++) {
        if (gamma * pla[i] / dla[i] < eta2 * (ula[i] * ula[i]) ||
            fmx(fmx(fabs(cm[i - 1]), fabs(c0[i - 1])), fabs(cp[i - 1])) < 1.0e-3 ||
            S2(dls, i, j) / dla[i] > 5.0) {
          for (ic = 1; ic <= ncolor; ic++)
            S3(colls, i, j, ic) = S3(colls, i, j, ic) * dla[i] / S2(dls, i, j);
          S2(pls, i, j) = pla[i];
          S2(uls, i, j) = ula[i];
          S2(dls, i, j) = dla[i];
        }
        if (gamma * pra[i] / dra[i] < eta2 * (ura[i] * ura[i]) ||
            fmx(fmx(fabs(cm[i]), fabs(c0[i])), fabs(cp[i])) < 1.0e-3 ||
            S2(drs, i, j) / dra[i] > 5.0) {
          for (ic = 1; ic <= ncolor; ic++)
            S3(colrs, i, j, ic) = S3(colrs, i, j, ic) * dra[i] / S2(drs, i, j);
          S2(prs, i, j) = pra[i];
          S2(urs, i, j) = ura[i];
          S2(drs, i, j) = dra[i];
        }
      }
    }
    /* floors (:540-551) */
    for (i = i1; i <= i2 + 1; i++) {
      S2(pls, i, j) = fmx(S2(pls, i, j), TINY);
      S2(prs, i, j) = fmx(S2(prs, i, j), TINY);
      S2(dls, i, j) = fmx(S2(dls, i, j), TINY);
      S2(drs, i, j) = fmx(S2(drs, i, j), TINY);
    }
    for (ic = 1; ic <= ncolor; ic++)
      for (i = i1; i <= i2 + 1; i++) {
        S3(colls, i, j, ic) = fmx(S3(colls, i, j, ic), COLOR_FLOOR);
        S3(colrs, i, j, ic) = fmx(S3(colrs, i, j, ic), COLOR_FLOOR);
      }
    if (ipresfree == 1)
      for (i = i1; i <= i2 + 1; i++) {
        S2(dls, i, j) = dla[i];
        S2(drs, i, j) = dra[i];
      }
  }
  free(colw);
}

/* ------------------------------------------------------------------ */
/* twoshock.F:5-276 -- two-shock Riemann solver, Newton on p*.          */
void orc_twoshock(const R *dls, const R *drs, R *pls, R *prs, const R *uls,
                  const R *urs, const I *pidim, const I *pjdim, const I *pi1,
                  const I *pi2, const I *pj1, const I *pj2, const R *pdt,
                  const R *pgamma, const R *ppmin, const I *pipresfree,
                  R *pbar, R *ubar, const I *pgravity, const R *grslice,
                  const I *pidual, const R *peta1) {
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const R gamma = *pgamma, pmin = *ppmin;
  const R tolerance = 1.e-14; /* :97-100 (CONFIG_BFLOAT_8) */
  const I numiter = 8;        /* :106 */
  (void)pdt; (void)pgravity; (void)grslice; (void)pidual; (void)peta1; (void)jdim;
  I i, j, n;
  const R qa = (gamma + 1.0) / (2.0 * gamma);
  for (j = j1; j <= j2; j++) {
    if (*pipresfree == 1) {
      for (i = i1; i <= i2; i++) {
        S2(pbar, i, j) = pmin;
        S2(ubar, i, j) = 0.5 * (S2(uls, i, j) + S2(urs, i, j));
        S2(pls, i, j) = pmin;
        S2(prs, i, j) = pmin;
      }
      continue;
    }
    for (i = i1; i <= i2; i++) {
      /* initial guess (:156-170); the gravity == 99 branch is never taken */
      R cl = sqrt(gamma * S2(pls, i, j) * S2(dls, i, j));
      R cr = sqrt(gamma * S2(prs, i, j) * S2(drs, i, j));
      R ps = (cr * S2(pls, i, j) + cl * S2(prs, i, j) +
              cr * cl * (S2(uls, i, j) - S2(urs, i, j))) / (cr + cl);
      if (ps < pmin) ps = pmin;
      R old_ps = ps, delta_ps;
      int mask = 1;
      R zl = 0, zr = 0, ubl = 0, ubr = 0, dpdul = 0, dpdur = 0;
      /* Newton iterations (:184-220); values freeze once converged */
      for (n = 2; n <= numiter; n++) {
        if (mask > 0) {
          zl = cl * sqrt((1.0 + qa * (ps / S2(pls, i, j) - 1.0)));
          zr = cr * sqrt((1.0 + qa * (ps / S2(prs, i, j) - 1.0)));
          ubl = S2(uls, i, j) - (ps - S2(pls, i, j)) / zl;
          ubr = S2(urs, i, j) + (ps - S2(prs, i, j)) / zr;
          dpdul = -(4.0 * (zl * (zl * zl)) / S2(dls, i, j) /
                    (4.0 * (zl * zl) / S2(dls, i, j) - (gamma + 1.0) * (ps - S2(pls, i, j))));
          dpdur = 4.0 * (zr * (zr * zr)) / S2(drs, i, j) /
                  (4.0 * (zr * zr) / S2(drs, i, j) - (gamma + 1.0) * (ps - S2(prs, i, j)));
          ps = ps + (ubr - ubl) * dpdur * dpdul / (dpdur - dpdul);
          if (ps < pmin) ps = pmin;
          delta_ps = ps - old_ps;
          old_ps = ps;
          if (fabs(delta_ps / ps) < tolerance) mask = 0;
        }
      }
      /* resolved state (:249-254) */
      if (ps < pmin) ps = fmn(S2(pls, i, j), S2(prs, i, j));
      S2(pbar, i, j) = ps;
      S2(ubar, i, j) = ubl + (ubr - ubl) * dpdur / (dpdur - dpdul);
    }
  }
}

/* ------------------------------------------------------------------ */
/* flux_hll.F:6-415 -- HLL fluxes (here only the one-cell fallback).
 * ext_ges: when nonzero, sl/gesl at interface i2+2 are computed so that
 * ges(i2+1) is defined (the reference reads them uninitialised).       */
static void orc_flux_hll_impl(const R *dslice, const R *eslice, const R *geslice,
                              const R *uslice, const R *vslice, const R *wslice,
                              const R *dx, const R *diffcoef, I idim, I jdim, I i1,
                              I i2, I j1, I j2, R dt, R gamma, I idiff, I idual,
                              const R *dls, const R *drs, const R *pls, const R *prs,
                              const R *uls, const R *urs, const R *vls, const R *vrs,
                              const R *wls, const R *wrs, const R *gels, const R *gers,
                              R *df, R *uf, R *vf, R *wf, R *ef, R *gef, R *ges,
                              I ncolor, const R *colslice, const R *colls,
                              const R *colrs, R *colf, int ext_ges) {
  const R gamma1 = gamma - 1.0;
  const R gamma1i = 1.0 / gamma1;
  I i, j, n;
  (void)jdim;
  const I ie = ext_ges ? i2 + 2 : i2 + 1; /* upper interface for sl/gesl */
  R sl[IJKN + 4], sr[IJKN + 4], gesl[IJKN + 4], gesr[IJKN + 4], bm0[IJKN + 4], bp0[IJKN + 4];
  R el[IJKN + 4], er[IJKN + 4], char1[IJKN + 4], char2[IJKN + 4], csv[IJKN + 4];
  R bm[IJKN + 4], bp[IJKN + 4];
  for (j = j1; j <= j2; j++) {
    /* Roe averages (:107-137) */
    for (i = i1; i <= ie; i++) {
      R sqrtdl = sqrt(S2(dls, i, j));
      R sqrtdr = sqrt(S2(drs, i, j));
      R isdlpdr = 1.0 / (sqrtdl + sqrtdr);
      R vroe1 = (sqrtdl * S2(uls, i, j) + sqrtdr * S2(urs, i, j)) * isdlpdr;
      R vroe2 = (sqrtdl * S2(vls, i, j) + sqrtdr * S2(vrs, i, j)) * isdlpdr;
      R vroe3 = (sqrtdl * S2(wls, i, j) + sqrtdr * S2(wrs, i, j)) * isdlpdr;
      R v2 = vroe1 * vroe1 + vroe2 * vroe2 + vroe3 * vroe3;
      el[i] = gamma1i * S2(pls, i, j) + 0.5 * S2(dls, i, j) *
              (S2(uls, i, j) * S2(uls, i, j) + S2(vls, i, j) * S2(vls, i, j) +
               S2(wls, i, j) * S2(wls, i, j));
      er[i] = gamma1i * S2(prs, i, j) + 0.5 * S2(drs, i, j) *
              (S2(urs, i, j) * S2(urs, i, j) + S2(vrs, i, j) * S2(vrs, i, j) +
               S2(wrs, i, j) * S2(wrs, i, j));
      R hroe = ((el[i] + S2(pls, i, j)) / sqrtdl + (er[i] + S2(prs, i, j)) / sqrtdr) * isdlpdr;
      csv[i] = sqrt(gamma1 * fmx((hroe - 0.5 * v2), TINY));
      char1[i] = vroe1 - csv[i];
      char2[i] = vroe1 + csv[i];
    }
    /* Davis wave speeds (:141-156) */
    for (i = i1; i <= ie; i++) {
      R csl0 = sqrt(gamma * S2(pls, i, j) / S2(dls, i, j));
      R csr0 = sqrt(gamma * S2(prs, i, j) / S2(drs, i, j));
      R csl = fmn(S2(uls, i, j) - csl0, char1[i]);
      R csr = fmx(S2(urs, i, j) + csr0, char2[i]);
      bm[i] = fmn(csl, 0.0);
      bp[i] = fmx(csr, 0.0);
      bm0[i] = S2(uls, i, j) - bm[i];
      bp0[i] = S2(urs, i, j) - bp[i];
    }
    /* HLL weights (:158-162) */
    for (i = i1; i <= ie; i++) {
      R q1 = (bp[i] + bm[i]) / (bp[i] - bm[i]);
      sl[i] = 0.5 * (1.0 + q1);
      sr[i] = 0.5 * (1.0 - q1);
    }
    for (i = i1; i <= i2 + 1; i++) {
      R diffd = 0, diffuu = 0, diffuv = 0, diffuw = 0, diffue = 0, diffuge = 0;
      if (idiff != 0) { /* :166-195 */
        R dc = S2(diffcoef, i, j);
        diffd = dc * (S2(dslice, i - 1, j) - S2(dslice, i, j));
        diffuu = dc * (S2(dslice, i - 1, j) * S2(uslice, i - 1, j) - S2(dslice, i, j) * S2(uslice, i, j));
        diffuv = dc * (S2(dslice, i - 1, j) * S2(vslice, i - 1, j) - S2(dslice, i, j) * S2(vslice, i, j));
        diffuw = dc * (S2(dslice, i - 1, j) * S2(wslice, i - 1, j) - S2(dslice, i, j) * S2(wslice, i, j));
        diffue = dc * (S2(dslice, i - 1, j) * S2(eslice, i - 1, j) - S2(dslice, i, j) * S2(eslice, i, j));
        if (idual == 1)
          diffuge = dc * (S2(dslice, i - 1, j) * S2(geslice, i - 1, j) - S2(dslice, i, j) * S2(geslice, i, j));
      }
      /* left/right fluxes (:202-226) */
      R dubl = S2(dls, i, j) * S2(uls, i, j);
      R dubr = S2(drs, i, j) * S2(urs, i, j);
      R dfl = S2(dls, i, j) * bm0[i];
      R dfr = S2(drs, i, j) * bp0[i];
      R ufl = dubl * bm0[i] + S2(pls, i, j);
      R ufr = dubr * bp0[i] + S2(prs, i, j);
      R vfl = S2(dls, i, j) * S2(vls, i, j) * bm0[i];
      R vfr = S2(drs, i, j) * S2(vrs, i, j) * bp0[i];
      R wfl = S2(dls, i, j) * S2(wls, i, j) * bm0[i];
      R wfr = S2(drs, i, j) * S2(wrs, i, j) * bp0[i];
      R efl = el[i] * bm0[i] + S2(pls, i, j) * S2(uls, i, j);
      R efr = er[i] * bp0[i] + S2(prs, i, j) * S2(urs, i, j);
      /* HLL flux plus diffusion (:252-270) */
      S2(df, i, j) = sl[i] * dfl + sr[i] * dfr + diffd;
      S2(uf, i, j) = sl[i] * ufl + sr[i] * ufr + diffuu;
      S2(vf, i, j) = sl[i] * vfl + sr[i] * vfr + diffuv;
      S2(wf, i, j) = sl[i] * wfl + sr[i] * wfr + diffuw;
      S2(ef, i, j) = sl[i] * efl + sr[i] * efr + diffue;
      if (idual == 1) {
        R gefl = bm0[i] * S2(gels, i, j) * S2(dls, i, j);
        R gefr = bp0[i] * S2(gers, i, j) * S2(drs, i, j);
        S2(gef, i, j) = sl[i] * gefl + sr[i] * gefr + diffuge;
      }
      for (n = 1; n <= ncolor; n++) {
        R colfl = bm0[i] * S3(colls, i, j, n);
        R colfr = bp0[i] * S3(colrs, i, j, n);
        S3(colf, i, j, n) = sl[i] * colfl + sr[i] * colfr;
      }
    }
    if (idual == 1) {
      for (i = i1; i <= ie; i++) { gesl[i] = bm0[i]; gesr[i] = bp0[i]; }
      /* gas-energy source term (:297-308) */
      for (i = i1; i <= i2 + 1; i++) {
        R pcent = fmx((gamma - 1.0) * S2(geslice, i, j) * S2(dslice, i, j), TINY);
        R qc = dt / F1(dx, i);
        S2(ges, i, j) = qc * pcent *
                        (sl[i] * gesl[i] + sr[i] * gesr[i] - sl[i + 1] * gesl[i + 1] -
                         sr[i + 1] * gesr[i + 1]);
      }
    }
    /* absorb dt/dx (:313-332) */
    for (i = i1; i <= i2 + 1; i++) {
      R qc = dt / F1(dx, i);
      S2(df, i, j) = qc * S2(df, i, j);
      S2(uf, i, j) = qc * S2(uf, i, j);
      S2(vf, i, j) = qc * S2(vf, i, j);
      S2(wf, i, j) = qc * S2(wf, i, j);
      S2(ef, i, j) = qc * S2(ef, i, j);
    }
    if (idual == 1)
      for (i = i1; i <= i2 + 1; i++) S2(gef, i, j) = dt / F1(dx, i) * S2(gef, i, j);
    for (n = 1; n <= ncolor; n++)
      for (i = i1; i <= i2 + 1; i++) S3(colf, i, j, n) = dt * S3(colf, i, j, n);
    /* negative density check -> abort (:336-384) */
    for (i = i1; i <= i2; i++)
      if (S2(dslice, i, j) + S2(df, i, j) - S2(df, i + 1, j) <= 0.0) orc_error_code = 5;
  }
}

void orc_flux_hll(const R *dslice, const R *eslice, const R *geslice,
                  const R *uslice, const R *vslice, const R *wslice, const R *dx,
                  const R *diffcoef, const I *idim, const I *jdim, const I *i1,
                  const I *i2, const I *j1, const I *j2, const R *dt,
                  const R *gamma, const I *idiff, const I *idual, const R *eta1,
                  const I *ifallback, const R *dls, const R *drs, const R *pls,
                  const R *prs, const R *uls, const R *urs, const R *vls,
                  const R *vrs, const R *wls, const R *wrs, const R *gels,
                  const R *gers, R *df, R *uf, R *vf, R *wf, R *ef, R *gef,
                  R *ges, const I *ncolor, const R *colslice, const R *colls,
                  const R *colrs, R *colf) {
  (void)eta1; (void)ifallback;
  /* dummy-argument order of flux_hll.F:6-12: df, uf, vf, wf, ef, gef, ges */
  orc_flux_hll_impl(dslice, eslice, geslice, uslice, vslice, wslice, dx, diffcoef,
                    *idim, *jdim, *i1, *i2, *j1, *j2, *dt, *gamma, *idiff, *idual,
                    dls, drs, pls, prs, uls, urs, vls, vrs, wls, wrs, gels, gers, df,
                    uf, vf, wf, ef, gef, ges, *ncolor, colslice, colls, colrs, colf, 0);
}

/* ------------------------------------------------------------------ */
/* flux_hllc.F:6-421 -- HLLC fluxes over the slice (RiemannSolver HLLC):
 * Roe averages and Davis wave speeds as in flux_hll, the contact speed cw
 * and pressure cp, HLLC weights sl / sr / sm chosen by the sign of cw, the
 * left/right fluxes with dfl = dubl - bm*dls, the contact terms sm*cp (x
 * momentum) and sm*cp*cw (energy), the gas-energy source, dt/dx, then the
 * density check with the per-cell flux_hll fallback in cell order. */
void orc_flux_hllc(const R *dslice, const R *eslice, const R *geslice,
                   const R *uslice, const R *vslice, const R *wslice, const R *dx,
                   const R *diffcoef, const I *pidim, const I *pjdim, const I *pi1,
                   const I *pi2, const I *pj1, const I *pj2, const R *pdt,
                   const R *pgamma, const I *pidiff, const I *pidual, const R *eta1,
                   const I *pifallback, const R *dls, const R *drs, const R *pls,
                   const R *prs, const R *uls, const R *urs, const R *vls,
                   const R *vrs, const R *wls, const R *wrs, const R *gels,
                   const R *gers, R *df, R *uf, R *vf, R *wf, R *ef, R *gef,
                   R *ges, const I *pncolor, const R *colslice, const R *colls,
                   const R *colrs, R *colf) {
  (void)eta1;
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const I idiff = *pidiff, idual = *pidual, ncolor = *pncolor;
  const R dt = *pdt, gamma = *pgamma;
  const R gamma1 = gamma - 1.0;
  const R gamma1i = 1.0 / gamma1;
  I i, j, n;
  (void)jdim;
  R el[IJKN + 4], er[IJKN + 4], char1[IJKN + 4], char2[IJKN + 4];
  R csl[IJKN + 4], csr[IJKN + 4], bm[IJKN + 4], bp[IJKN + 4];
  R cw[IJKN + 4], cp[IJKN + 4], sl[IJKN + 4], sr[IJKN + 4], sm[IJKN + 4];
  R gesl[IJKN + 4], gesr[IJKN + 4];
  for (j = j1; j <= j2; j++) {
    /* Roe averages, characteristics (:104-136) */
    for (i = i1; i <= i2 + 1; i++) {
      R sqrtdl = sqrt(S2(dls, i, j));
      R sqrtdr = sqrt(S2(drs, i, j));
      R isdlpdr = 1.0 / (sqrtdl + sqrtdr);
      R vroe1 = (sqrtdl * S2(uls, i, j) + sqrtdr * S2(urs, i, j)) * isdlpdr;
      R vroe2 = (sqrtdl * S2(vls, i, j) + sqrtdr * S2(vrs, i, j)) * isdlpdr;
      R vroe3 = (sqrtdl * S2(wls, i, j) + sqrtdr * S2(wrs, i, j)) * isdlpdr;
      R v2 = vroe1 * vroe1 + vroe2 * vroe2 + vroe3 * vroe3;
      el[i] = gamma1i * S2(pls, i, j) + 0.5 * S2(dls, i, j) *
              (S2(uls, i, j) * S2(uls, i, j) + S2(vls, i, j) * S2(vls, i, j) +
               S2(wls, i, j) * S2(wls, i, j));
      er[i] = gamma1i * S2(prs, i, j) + 0.5 * S2(drs, i, j) *
              (S2(urs, i, j) * S2(urs, i, j) + S2(vrs, i, j) * S2(vrs, i, j) +
               S2(wrs, i, j) * S2(wrs, i, j));
      R hroe = ((el[i] + S2(pls, i, j)) / sqrtdl + (er[i] + S2(prs, i, j)) / sqrtdr) * isdlpdr;
      R cs = sqrt(gamma1 * fmx((hroe - 0.5 * v2), TINY));
      char1[i] = vroe1 - cs;
      char2[i] = vroe1 + cs;
    }
    /* min/max wave speeds (:140-153) */
    for (i = i1; i <= i2 + 1; i++) {
      R csl0 = sqrt(gamma * S2(pls, i, j) / S2(dls, i, j));
      R csr0 = sqrt(gamma * S2(prs, i, j) / S2(drs, i, j));
      csl[i] = fmn(S2(uls, i, j) - csl0, char1[i]);
      csr[i] = fmx(S2(urs, i, j) + csr0, char2[i]);
      bm[i] = fmn(csl[i], 0.0);
      bp[i] = fmx(csr[i], 0.0);
    }
    /* contact speed and pressure, HLLC weights (:158-180) */
    for (i = i1; i <= i2 + 1; i++) {
      R tl = S2(pls, i, j) - (csl[i] - S2(uls, i, j)) * S2(dls, i, j) * S2(uls, i, j);
      R tr = S2(prs, i, j) - (csr[i] - S2(urs, i, j)) * S2(drs, i, j) * S2(urs, i, j);
      R dl = S2(dls, i, j) * (csl[i] - S2(uls, i, j));
      R dr = -S2(drs, i, j) * (csr[i] - S2(urs, i, j));
      R q1 = 1.0 / (dl + dr);
      cw[i] = (tr - tl) * q1;
      cp[i] = (dl * tr + dr * tl) * q1;
      if (cw[i] >= 0.0) {
        sl[i] = cw[i] / (cw[i] - bm[i]);
        sr[i] = 0.0;
        sm[i] = -bm[i] / (cw[i] - bm[i]);
      } else {
        sl[i] = 0.0;
        sr[i] = -cw[i] / (bp[i] - cw[i]);
        sm[i] = bp[i] / (bp[i] - cw[i]);
      }
      cp[i] = fmx(cp[i], 0.0);
    }
    for (i = i1; i <= i2 + 1; i++) {
      R diffd = 0, diffuu = 0, diffuv = 0, diffuw = 0, diffue = 0, diffuge = 0;
      if (idiff != 0) { /* :185-215 */
        R dc = S2(diffcoef, i, j);
        diffd = dc * (S2(dslice, i - 1, j) - S2(dslice, i, j));
        diffuu = dc * (S2(dslice, i - 1, j) * S2(uslice, i - 1, j) - S2(dslice, i, j) * S2(uslice, i, j));
        diffuv = dc * (S2(dslice, i - 1, j) * S2(vslice, i - 1, j) - S2(dslice, i, j) * S2(vslice, i, j));
        diffuw = dc * (S2(dslice, i - 1, j) * S2(wslice, i - 1, j) - S2(dslice, i, j) * S2(wslice, i, j));
        diffue = dc * (S2(dslice, i - 1, j) * S2(eslice, i - 1, j) - S2(dslice, i, j) * S2(eslice, i, j));
        if (idual == 1)
          diffuge = dc * (S2(dslice, i - 1, j) * S2(geslice, i - 1, j) - S2(dslice, i, j) * S2(geslice, i, j));
      }
      /* left/right fluxes (:222-244) */
      R dubl = S2(dls, i, j) * S2(uls, i, j);
      R dubr = S2(drs, i, j) * S2(urs, i, j);
      R dfl = dubl - bm[i] * S2(dls, i, j);
      R dfr = dubr - bp[i] * S2(drs, i, j);
      R ufl = dubl * (S2(uls, i, j) - bm[i]) + S2(pls, i, j);
      R ufr = dubr * (S2(urs, i, j) - bp[i]) + S2(prs, i, j);
      R vfl = S2(dls, i, j) * S2(vls, i, j) * (S2(uls, i, j) - bm[i]);
      R vfr = S2(drs, i, j) * S2(vrs, i, j) * (S2(urs, i, j) - bp[i]);
      R wfl = S2(dls, i, j) * S2(wls, i, j) * (S2(uls, i, j) - bm[i]);
      R wfr = S2(drs, i, j) * S2(wrs, i, j) * (S2(urs, i, j) - bp[i]);
      R efl = el[i] * (S2(uls, i, j) - bm[i]) + S2(pls, i, j) * S2(uls, i, j);
      R efr = er[i] * (S2(urs, i, j) - bp[i]) + S2(prs, i, j) * S2(urs, i, j);
      /* HLLC flux plus diffusion (:272-290), contact terms (:299-302) */
      S2(df, i, j) = sl[i] * dfl + sr[i] * dfr + diffd;
      S2(uf, i, j) = sl[i] * ufl + sr[i] * ufr + diffuu;
      S2(vf, i, j) = sl[i] * vfl + sr[i] * vfr + diffuv;
      S2(wf, i, j) = sl[i] * wfl + sr[i] * wfr + diffuw;
      S2(ef, i, j) = sl[i] * efl + sr[i] * efr + diffue;
      if (idual == 1) {
        R gefl = (S2(uls, i, j) - bm[i]) * S2(gels, i, j) * S2(dls, i, j);
        R gefr = (S2(urs, i, j) - bp[i]) * S2(gers, i, j) * S2(drs, i, j);
        gesl[i] = S2(uls, i, j) - bm[i];
        gesr[i] = S2(urs, i, j) - bp[i];
        S2(gef, i, j) = sl[i] * gefl + sr[i] * gefr + diffuge;
      }
      for (n = 1; n <= ncolor; n++) {
        R colfl = (S2(uls, i, j) - bm[i]) * S3(colls, i, j, n);
        R colfr = (S2(urs, i, j) - bp[i]) * S3(colrs, i, j, n);
        S3(colf, i, j, n) = sl[i] * colfl + sr[i] * colfr;
      }
      S2(uf, i, j) = S2(uf, i, j) + sm[i] * cp[i];
      S2(ef, i, j) = S2(ef, i, j) + sm[i] * cp[i] * cw[i];
    }
    /* gas-energy source (:306-316); the i = i2+1 value reads the locals
       at i2+2, which the reference never sets and euler never uses */
    if (idual == 1) {
      sl[i2 + 2] = sr[i2 + 2] = gesl[i2 + 2] = gesr[i2 + 2] = 0.0;
      for (i = i1; i <= i2 + 1; i++) {
        R pcent = fmx((gamma - 1.0) * S2(geslice, i, j) * S2(dslice, i, j), TINY);
        R qc = dt / F1(dx, i);
        S2(ges, i, j) = qc * pcent *
                        (sl[i] * gesl[i] + sr[i] * gesr[i] - sl[i + 1] * gesl[i + 1] -
                         sr[i + 1] * gesr[i + 1]);
      }
    }
    /* absorb dt/dx (:321-340) */
    for (i = i1; i <= i2 + 1; i++) {
      R qc = dt / F1(dx, i);
      S2(df, i, j) = qc * S2(df, i, j);
      S2(uf, i, j) = qc * S2(uf, i, j);
      S2(vf, i, j) = qc * S2(vf, i, j);
      S2(wf, i, j) = qc * S2(wf, i, j);
      S2(ef, i, j) = qc * S2(ef, i, j);
    }
    if (idual == 1)
      for (i = i1; i <= i2 + 1; i++) S2(gef, i, j) = dt / F1(dx, i) * S2(gef, i, j);
    for (n = 1; n <= ncolor; n++)
      for (i = i1; i <= i2 + 1; i++) S3(colf, i, j, n) = dt * S3(colf, i, j, n);
    /* density check and per-cell flux_hll fallback, in cell order (:344-391) */
    for (i = i1; i <= i2; i++) {
      if (S2(dslice, i, j) + S2(df, i, j) - S2(df, i + 1, j) <= 0.0) {
        if (*pifallback == 0) {
          orc_error_code = 9;
        } else {
          orc_fallback_count++;
          orc_flux_hll_impl(dslice, eslice, geslice, uslice, vslice, wslice, dx,
                            diffcoef, idim, jdim, i, i, j, j, dt, gamma, idiff,
                            idual, dls, drs, pls, prs, uls, urs, vls, vrs, wls,
                            wrs, gels, gers, df, uf, vf, wf, ef, gef, ges, ncolor,
                            colslice, colls, colrs, colf, 1);
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* flux_twoshock.F:6-490 -- Godunov fluxes from the two-shock solution. */
void orc_flux_twoshock(const R *dslice, const R *eslice, const R *geslice,
                       const R *uslice, const R *vslice, const R *wslice,
                       const R *dx, const R *diffcoef, const I *pidim,
                       const I *pjdim, const I *pi1, const I *pi2, const I *pj1,
                       const I *pj2, const R *pdt, const R *pgamma, const I *pidiff,
                       const I *pidual, const R *peta1, const I *pifallback,
                       const R *dls, const R *drs, const R *pls, const R *prs,
                       const R *gels, const R *gers, const R *uls, const R *urs,
                       const R *vls, const R *vrs, const R *wls, const R *wrs,
                       const R *pbar, const R *ubar, R *df, R *ef, R *uf, R *vf,
                       R *wf, R *gef, R *ges, const I *pncolor, const R *colslice,
                       const R *colls, const R *colrs, R *colf) {
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const I idiff = *pidiff, idual = *pidual, ncolor = *pncolor;
  const R dt = *pdt, gamma = *pgamma;
  (void)peta1;
  I i, j, n;
  const R qa = (gamma + 1.0) / (2.0 * gamma);
  R ub[IJKN + 4];
  for (j = j1; j <= j2; j++) {
    for (i = i1; i <= i2 + 1; i++) {
      /* upwind state (:84-109) */
      R sn = sgn1(-S2(ubar, i, j));
      R u0, p0, d0;
      if (sn < 0.0) { u0 = S2(uls, i, j); p0 = S2(pls, i, j); d0 = S2(dls, i, j); }
      else { u0 = S2(urs, i, j); p0 = S2(prs, i, j); d0 = S2(drs, i, j); }
      R c0 = sqrt(fmx(gamma * p0 / d0, TINY));
      R z0 = c0 * d0 * sqrt(fmx(1.0 + qa * (S2(pbar, i, j) / p0 - 1.0), TINY));
      /* post-wave density and sound speed (:128-132) */
      R dbar = 1.0 / (1.0 / d0 - (S2(pbar, i, j) - p0) / fmx(z0 * z0, TINY));
      R cbar = sqrt(fmx(gamma * S2(pbar, i, j) / dbar, TINY));
      /* wave speeds (:137-146) */
      R l0, lbar;
      if (S2(pbar, i, j) < p0) {
        l0 = u0 * sn + c0;
        lbar = sn * S2(ubar, i, j) + cbar;
      } else {
        l0 = u0 * sn + z0 / d0;
        lbar = l0;
      }
      /* RAREFACTION2 linear fan (:185-194) */
      R frac = l0 - lbar;
      if (frac < TINY) frac = TINY;
      frac = (0.0 - lbar) / frac;
      frac = fmn(fmx(frac, 0.0), 1.0);
      R pb = p0 * frac + S2(pbar, i, j) * (1.0 - frac);
      R db = d0 * frac + dbar * (1.0 - frac);
      R ubv = u0 * frac + S2(ubar, i, j) * (1.0 - frac);
      /* sampling (:201-218) */
      if (lbar >= 0.0) { pb = S2(pbar, i, j); db = dbar; ubv = S2(ubar, i, j); }
      if (l0 < 0.0) { pb = p0; db = d0; ubv = u0; }
      ub[i] = ubv;
      /* advected quantities (:237-259) */
      R vb, wb, geb;
      if (ubv > 0.0) { vb = S2(vls, i, j); wb = S2(wls, i, j); geb = S2(gels, i, j); }
      else { vb = S2(vrs, i, j); wb = S2(wrs, i, j); geb = S2(gers, i, j); }
      /* total specific energy (:283-287) */
      R eb = pb / ((gamma - 1.0) * db) + 0.5 * (ubv * ubv + vb * vb + wb * wb);
      R upb, dub, duub, duvb, duwb, dueb, dugeb = 0.0;
      if (idiff != 0) { /* :291-322 */
        R dc = S2(diffcoef, i, j);
        upb = pb * ubv;
        dub = ubv * db;
        duub = dub * ubv + dc * (S2(dslice, i - 1, j) * S2(uslice, i - 1, j) - S2(dslice, i, j) * S2(uslice, i, j));
        duvb = dub * vb + dc * (S2(dslice, i - 1, j) * S2(vslice, i - 1, j) - S2(dslice, i, j) * S2(vslice, i, j));
        duwb = dub * wb + dc * (S2(dslice, i - 1, j) * S2(wslice, i - 1, j) - S2(dslice, i, j) * S2(wslice, i, j));
        dueb = dub * eb + dc * (S2(dslice, i - 1, j) * S2(eslice, i - 1, j) - S2(dslice, i, j) * S2(eslice, i, j));
        dub = dub + dc * (S2(dslice, i - 1, j) - S2(dslice, i, j));
        if (idual == 1)
          dugeb = dub * geb + dc * (S2(dslice, i - 1, j) * S2(geslice, i - 1, j) - S2(dslice, i, j) * S2(geslice, i, j));
      } else { /* :326-340 */
        upb = pb * ubv;
        dub = ubv * db;
        duub = dub * ubv;
        duvb = dub * vb;
        duwb = dub * wb;
        dueb = dub * eb;
        if (idual == 1) dugeb = dub * geb;
      }
      /* fluxes times dt/dx (:365-372) */
      R qc = dt / F1(dx, i);
      S2(df, i, j) = qc * dub;
      S2(ef, i, j) = qc * (dueb + upb);
      S2(uf, i, j) = qc * (duub + pb);
      S2(vf, i, j) = qc * duvb;
      S2(wf, i, j) = qc * duwb;
      /* colour fluxes dt*u_b*c_b (:247-259, :375-380) */
      for (n = 1; n <= ncolor; n++) {
        R colb;
        if (ubv > 0.0) colb = S3(colls, i, j, n) * db / S2(dls, i, j);
        else colb = S3(colrs, i, j, n) * db / S2(drs, i, j);
        S3(colf, i, j, n) = dt * ubv * colb;
      }
      if (idual == 1) S2(gef, i, j) = qc * dugeb;
    }
    /* gas-energy source term (:386-396); the i = i2+1 value reads ub(i2+2),
       which the reference never sets and euler never uses; it is zero here */
    if (idual == 1) {
      ub[i2 + 2] = 0.0;
      for (i = i1; i <= i2 + 1; i++) {
        R pcent = fmx((gamma - 1.0) * S2(geslice, i, j) * S2(dslice, i, j), TINY);
        R qc = dt / F1(dx, i);
        S2(ges, i, j) = qc * pcent * (ub[i] - ub[i + 1]);
      }
    }
    /* negative density/energy check and per-cell HLL fallback (:401-456) */
    for (i = i1; i <= i2; i++) {
      if (S2(dslice, i, j) + (S2(df, i, j) - S2(df, i + 1, j)) <= 0.0 ||
          S2(eslice, i, j) < 0.0) {
        if (*pifallback == 0) {
          orc_error_code = 6;
        } else {
          orc_fallback_count++;
          orc_flux_hll_impl(dslice, eslice, geslice, uslice, vslice, wslice, dx,
                            diffcoef, idim, jdim, i, i, j, j, dt, gamma, idiff,
                            idual, dls, drs, pls, prs, uls, urs, vls, vrs, wls,
                            wrs, gels, gers, df, uf, vf, wf, ef, gef, ges, ncolor,
                            colslice, colls, colrs, colf, 1);
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* euler.F:7-284 -- conservative update of the slice.                   */
void orc_euler(R *dslice, R *eslice, const R *grslice, R *geslice, R *uslice,
               R *vslice, R *wslice, const R *dx, const R *diffcoef,
               const I *pidim, const I *pjdim, const I *pi1, const I *pi2,
               const I *pj1, const I *pj2, const R *pdt, const R *pgamma,
               const I *pidiff, const I *pgravity, const I *pidual,
               const R *peta1, const R *peta2, R *df, R *ef, R *uf, R *vf,
               R *wf, R *gef, R *ges, const I *pncolor, R *colslice, R *colf,
               const R *pdfloor) {
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const I gravity = *pgravity, idual = *pidual, ncolor = *pncolor;
  const R dt = *pdt, dfloor = *pdfloor;
  const R min_color = 1e-5 * TINY; /* :104-105 */
  (void)diffcoef; (void)pidiff; (void)peta1; (void)peta2; (void)pgamma;
  I i, j, n;
  R dnu[IJKN + 2], dnuinv[IJKN + 2], uold[IJKN + 2];
  for (j = j1; j <= j2; j++) {
    for (i = i1; i <= i2; i++) { /* :126-169 */
      dnu[i] = S2(dslice, i, j) + (S2(df, i, j) - S2(df, i + 1, j));
      dnuinv[i] = 1.0 / dnu[i];
      if (dfloor > 0.0) {
        dnu[i] = fmx(dnu[i], dfloor);
        dnuinv[i] = 1.0 / dnu[i];
      }
      uold[i] = S2(uslice, i, j);
      S2(uslice, i, j) = (S2(uslice, i, j) * S2(dslice, i, j) + (S2(uf, i, j) - S2(uf, i + 1, j))) * dnuinv[i];
      S2(vslice, i, j) = (S2(vslice, i, j) * S2(dslice, i, j) + (S2(vf, i, j) - S2(vf, i + 1, j))) * dnuinv[i];
      S2(wslice, i, j) = (S2(wslice, i, j) * S2(dslice, i, j) + (S2(wf, i, j) - S2(wf, i + 1, j))) * dnuinv[i];
      S2(eslice, i, j) = fmx(0.1 * S2(eslice, i, j),
                             (S2(eslice, i, j) * S2(dslice, i, j) + (S2(ef, i, j) - S2(ef, i + 1, j))) * dnuinv[i]);
    }
    /* colours (:173-186) */
    for (n = 1; n <= ncolor; n++)
      for (i = i1; i <= i2 + 1; i++) S3(colf, i, j, n) = S3(colf, i, j, n) / F1(dx, i);
    for (n = 1; n <= ncolor; n++)
      for (i = i1; i <= i2; i++) {
        S3(colslice, i, j, n) = S3(colslice, i, j, n) + (S3(colf, i, j, n) - S3(colf, i + 1, j, n));
        S3(colslice, i, j, n) = fmx(S3(colslice, i, j, n), min_color);
      }
    /* gas energy with pdV source (:195-217) */
    if (idual == 1)
      for (i = i1; i <= i2; i++) {
        S2(geslice, i, j) = fmx((S2(geslice, i, j) * S2(dslice, i, j) +
                                 (S2(gef, i, j) - S2(gef, i + 1, j)) + S2(ges, i, j)) * dnuinv[i],
                                0.5 * S2(geslice, i, j));
        if (S2(geslice, i, j) < 0.0) orc_error_code = 7;
      }
    /* gravity, GRAVITY_METHOD1 (:248-273) */
    if (gravity == 1)
      for (i = i1; i <= i2; i++) {
        S2(uslice, i, j) = S2(uslice, i, j) + dt * S2(grslice, i, j) * 0.5 *
                                                  (S2(dslice, i, j) * dnuinv[i] + 1.0);
        S2(eslice, i, j) = S2(eslice, i, j) + dt * S2(grslice, i, j) * 0.5 *
                                                  (S2(uslice, i, j) + uold[i] * S2(dslice, i, j) * dnuinv[i]);
        S2(eslice, i, j) = fmx(S2(eslice, i, j), TINY);
      }
    for (i = i1; i <= i2; i++) S2(dslice, i, j) = dnu[i]; /* :277-279 */
  }
}

/* ------------------------------------------------------------------ */
/* pgas2d_dual.F:5-122 -- dual-energy consistency and slice pressure.    */
void orc_pgas2d_dual(const R *dslice, R *eslice, R *geslice, R *pslice,
                     const R *uslice, const R *vslice, const R *wslice,
                     const R *peta1, const R *peta2, const I *pidim,
                     const I *pjdim, const I *pi1, const I *pi2, const I *pj1,
                     const I *pj2, const R *pgamma, const R *ppmin) {
  const I idim = *pidim, jdim = *pjdim, i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2;
  const R eta1 = *peta1, eta2 = *peta2, gamma = *pgamma, pmin = *ppmin;
  (void)jdim;
  I i, j;
  for (j = j1; j <= j2; j++)
    for (i = i1; i <= i2; i++) {
      R ke = 0.5 * (S2(uslice, i, j) * S2(uslice, i, j) + S2(vslice, i, j) * S2(vslice, i, j) +
                    S2(wslice, i, j) * S2(wslice, i, j));
      R ge1 = S2(eslice, i, j) - ke;
      I im1 = (i - 1) > i1 ? (i - 1) : i1;
      I ip1 = (i + 1) < i2 ? (i + 1) : i2;
      R demax = fmx(fmx(S2(dslice, i, j) * S2(eslice, i, j), S2(dslice, im1, j) * S2(eslice, im1, j)),
                    S2(dslice, ip1, j) * S2(eslice, ip1, j));
      if (ge1 * S2(dslice, i, j) / demax > eta2) S2(geslice, i, j) = ge1;
      R ge2;
      if (ge1 / S2(eslice, i, j) > eta1) ge2 = ge1;
      else ge2 = S2(geslice, i, j);
      ge2 = fmx(ge2, pmin / ((gamma - 1.0) * S2(dslice, i, j)));
      S2(eslice, i, j) = S2(eslice, i, j) - ge1 + ge2;
      S2(pslice, i, j) = (gamma - 1.0) * S2(dslice, i, j) * ge2;
    }
}

/* ------------------------------------------------------------------ */
/* calc_dt.F:6-181 -- Courant time step (PPM: Godunov formula, ihydro=0). */
void orc_calc_dt(const I *prank, const I *pidim, const I *pjdim, const I *pkdim,
                 const I *pi1, const I *pi2, const I *pj1, const I *pj2,
                 const I *pk1, const I *pk2, const I *pihydro, const R *pC2,
                 const R *dx, const R *dy, const R *dz, const R *pvgx,
                 const R *pvgy, const R *pvgz, const R *pgamma, const I *pipfree,
                 const R *paye, const R *d, const R *p, const R *u, const R *v,
                 const R *w, R *pdt, R *pdtviscous) {
  const I rank = *prank, idim = *pidim, jdim = *pjdim;
  const I i1 = *pi1, i2 = *pi2, j1 = *pj1, j2 = *pj2, k1 = *pk1, k2 = *pk2;
  const R vgx = *pvgx, vgy = *pvgy, vgz = *pvgz, gamma = *pgamma, aye = *paye;
  (void)pkdim; (void)pC2; (void)pdtviscous;
  if (*pihydro == 2) { orc_error_code = 8; return; } /* Zeus viscosity: not this path */
  R dt = 1.0e20; /* huge */
  I i, j, k;
#define D3(a, i, j, k) (a)[((i) - 1) + ((j) - 1) * idim + ((k) - 1) * idim * jdim]
  if (rank == 1) {
    for (i = i1 + 1; i <= i2 + 1; i++) {
      R cs = fmx(sqrt(gamma * D3(p, i, 1, 1) / D3(d, i, 1, 1)), TINY);
      if (*pipfree == 1) cs = TINY;
      dt = fmn(dt, F1(dx, i) * aye / (cs + fabs(D3(u, i, 1, 1) - vgx)));
    }
  }
  if (rank == 2) {
    for (j = j1 + 1; j <= j2 + 1; j++)
      for (i = i1 + 1; i <= i2 + 1; i++) {
        R cs = fmx(sqrt(gamma * D3(p, i, j, 1) / D3(d, i, j, 1)), TINY);
        if (*pipfree == 1) cs = TINY;
        dt = fmn(dt, aye / ((cs + fabs(D3(u, i, j, 1) - vgx)) / F1(dx, i) +
                            (cs + fabs(D3(v, i, j, 1) - vgy)) / F1(dy, j)));
      }
  }
  if (rank == 3) {
    for (k = k1 + 1; k <= k2 + 1; k++)
      for (j = j1 + 1; j <= j2 + 1; j++)
        for (i = i1 + 1; i <= i2 + 1; i++) {
          R cs = fmx(sqrt(gamma * D3(p, i, j, k) / D3(d, i, j, k)), TINY);
          if (*pipfree == 1) cs = TINY;
          R dt1 = aye / ((cs + fabs(D3(u, i, j, k) - vgx)) / F1(dx, i) +
                         (cs + fabs(D3(v, i, j, k) - vgy)) / F1(dy, j) +
                         (cs + fabs(D3(w, i, j, k) - vgz)) / F1(dz, k));
          dt = fmn(dt, dt1);
        }
  }
#undef D3
  *pdt = dt;
}

/* ------------------------------------------------------------------ */
/* expand_terms.F:5-170 -- comoving expansion terms over isize cells,
 * ENERGY_METHOD3 and VELOCITY_METHOD3 (the #defines at :59 and :137);
 * p is the time-centred pressure, unchanged by these methods.  The old
 * fields are not read by these methods; they are kept in the argument list
 * as the reference has them. */
void orc_expand_terms(const I *prank, const I *pisize, const I *pidual, const R *pcoef,
                      const I *pimethod, const R *pgamma, R *p, const R *d, R *e, R *ge,
                      R *u, R *v, R *w, const R *dold, const R *eold, const R *geold,
                      const R *uold, const R *vold, const R *wold, const I *picr, R *ecr,
                      const R *ecrold) {
  const I rank = *prank, isize = *pisize, idual = *pidual, imethod = *pimethod, icr = *picr;
  const R coef = *pcoef, gamma = *pgamma;
  I i;
  (void)dold; (void)eold; (void)geold; (void)uold; (void)vold; (void)wold; (void)ecrold;
  /* gas energy (:66-83) */
  if (idual == 1)
    for (i = 0; i < isize; i++) ge[i] = ge[i] * (1.0 - coef) / (1.0 + coef);
  /* cosmic-ray energy (:88-92) */
  if (icr > 0)
    for (i = 0; i < isize; i++) ecr[i] = ecr[i] * (2.0 - coef) / (2.0 + coef);
  /* total energy (:93-130) */
  if (imethod == 2) {
    if (icr > 0)
      for (i = 0; i < isize; i++) e[i] = e[i] * (1.0 - coef) / (1.0 + coef);
    else
      for (i = 0; i < isize; i++)
        e[i] = fmx(e[i] - coef * 6.0 * p[i] / (d[i] + dold[i]), 0.5 * e[i]);
  } else {
    for (i = 0; i < isize; i++) {
      e[i] = e[i] * (1.0 - coef) / (1.0 + coef);
      e[i] = fmx(e[i] - coef * (3.0 - 2.0 / (gamma - 1.0)) * p[i] / d[i], 0.5 * e[i]);
    }
  }
  /* velocities, semi-implicit (:158-166) */
  for (i = 0; i < isize; i++) {
    u[i] = u[i] * (1.0 - 0.5 * coef) / (1.0 + 0.5 * coef);
    if (rank > 1) v[i] = v[i] * (1.0 - 0.5 * coef) / (1.0 + 0.5 * coef);
    if (rank > 2) w[i] = w[i] * (1.0 - 0.5 * coef) / (1.0 + 0.5 * coef);
  }
}
