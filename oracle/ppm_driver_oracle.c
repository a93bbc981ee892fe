/*
 * ORACLE -- test infrastructure only.
 *
 * Slice driver restating grid::SolvePPM_DE (Grid_SolvePPM_DE.C:36-164),
 * grid::[xyz]EulerSweep (Grid_xEulerSweep.C:33-463, Grid_yEulerSweep.C,
 * Grid_zEulerSweep.C), the time == Time branch of grid::ComputePressure
 * (Grid_ComputePressure.C:96-166) and the calc_dt call of
 * grid::ComputeTimeStep (Grid_ComputeTimeStep.C:149-209).
 *
 * This file is compiled twice (oracle/Makefile):
 *   ORACLE_PORT  -> kernels are this project's C restatement
 *                   (ppm_kernels_oracle.c), exported symbols orc_*;
 *   ORACLE_REF   -> kernels are the reference Fortran compiled with flang
 *                   from /root/reference (calcdiss_, inteuler_, ...),
 *                   exported symbols ref_*  (oracle/_ref/, never committed).
 * The driver itself is one restatement shared by both builds, so a
 * difference between the two builds isolates the per-slice kernels.
 *
 * Semantics kept from the reference:
 *   - pressure computed once per call and never refreshed between sweeps
 *     (Grid_SolvePPM_DE.C:58-61; pgas2d_dual only rewrites the slice copy);
 *   - sweep order rotates with CycleNumber (Grid_SolvePPM_DE.C:53,86);
 *   - every slice of the full transverse extent is swept, ghosts included
 *     (Grid_SolvePPM_DE.C:91,110,129; Grid_xEulerSweep.C:172-178);
 *   - calcdiss reads the live 3-D transverse velocities, so with diffusion on
 *     the slices run in order, one after another (Grid_xEulerSweep.C:197-206).
 * Without diffusion the slices are independent and run on a small pthread
 * pool (one set of slice buffers per thread).  Reference-side Fortran uses
 * large automatic arrays, so pool threads get 64 MiB stacks.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/enzo_ppm.h"

typedef int64_t I;
typedef double R;

#if defined(ORACLE_PORT)
#define PFX(n) orc_##n
#define KER(n) orc_##n
extern _Thread_local I orc_error_code;
extern _Thread_local I orc_fallback_count;
#define ERRCODE orc_error_code
#define FALLBACKS orc_fallback_count
#elif defined(ORACLE_REF)
#define PFX(n) ref_##n
#define KER(n) n##_
/* f_error / f_warning (f_message.F:17-45) replaced by recorders: the
   reference aborts the process on f_error; here the driver reports it. */
static _Thread_local I ref_error_code = 0;
static _Thread_local I ref_fallback_count = 0;
void f_error_(const char *file, const I *line, long len) {
  (void)file; (void)len;
  ref_error_code = *line > 0 ? *line : 1;
}
void f_warning_(const char *file, const I *line, long len) {
  /* flux_twoshock.F:446 and flux_hllc.F:380 warn once per cell before
   * calling flux_hll (their other warnings are about colours) */
  char name[512];  /* a Fortran string: len characters, no terminator */
  const size_t n = len > 0 ? ((size_t)len < sizeof(name) ? (size_t)len : sizeof(name) - 1) : 0;
  memcpy(name, file, n);
  name[n] = 0;
  if ((strstr(name, "flux_twoshock") != NULL && *line == 446) ||
      (strstr(name, "flux_hllc") != NULL && *line == 380))
    ref_fallback_count++;
}
#define ERRCODE ref_error_code
#define FALLBACKS ref_fallback_count
#else
#error "define ORACLE_PORT or ORACLE_REF"
#endif

/* per-slice kernels, Fortran calling convention (euler_sweep.h:1-88) */
void KER(calcdiss)(const R *, const R *, const R *, const R *, const R *, const R *,
                   const R *, const R *, const R *, const I *, const I *, const I *,
                   const I *, const I *, const I *, const I *, const I *, const I *,
                   const I *, const I *, const I *, const I *, const R *, const R *,
                   const I *, const I *, R *, R *);
void KER(inteuler)(const R *, const R *, const I *, const R *, const R *, const R *,
                   const R *, const R *, const R *, const R *, const I *, const I *,
                   const I *, const I *, const I *, const I *, const I *, const R *,
                   const R *, const I *, const I *, const I *, const I *, const R *,
                   const R *, const I *, R *, R *, R *, R *, R *, R *, R *, R *, R *,
                   R *, R *, R *, const I *, const R *, R *, R *);
void KER(twoshock)(const R *, const R *, R *, R *, const R *, const R *, const I *,
                   const I *, const I *, const I *, const I *, const I *, const R *,
                   const R *, const R *, const I *, R *, R *, const I *, const R *,
                   const I *, const R *);
void KER(flux_twoshock)(const R *, const R *, const R *, const R *, const R *,
                        const R *, const R *, const R *, const I *, const I *,
                        const I *, const I *, const I *, const I *, const R *,
                        const R *, const I *, const I *, const R *, const I *,
                        const R *, const R *, const R *, const R *, const R *,
                        const R *, const R *, const R *, const R *, const R *,
                        const R *, const R *, const R *, const R *, R *, R *, R *,
                        R *, R *, R *, R *, const I *, const R *, const R *,
                        const R *, R *);
void KER(flux_hll)(const R *, const R *, const R *, const R *, const R *, const R *,
                   const R *, const R *, const I *, const I *, const I *, const I *,
                   const I *, const I *, const R *, const R *, const I *, const I *,
                   const R *, const I *, const R *, const R *, const R *, const R *,
                   const R *, const R *, const R *, const R *, const R *, const R *,
                   const R *, const R *, R *, R *, R *, R *, R *, R *, R *, const I *,
                   const R *, const R *, const R *, R *);
void KER(flux_hllc)(const R *, const R *, const R *, const R *, const R *, const R *,
                    const R *, const R *, const I *, const I *, const I *, const I *,
                    const I *, const I *, const R *, const R *, const I *, const I *,
                    const R *, const I *, const R *, const R *, const R *, const R *,
                    const R *, const R *, const R *, const R *, const R *, const R *,
                    const R *, const R *, R *, R *, R *, R *, R *, R *, R *, const I *,
                    const R *, const R *, const R *, R *);
void KER(euler)(R *, R *, const R *, R *, R *, R *, R *, const R *, const R *,
                const I *, const I *, const I *, const I *, const I *, const I *,
                const R *, const R *, const I *, const I *, const I *, const R *,
                const R *, R *, R *, R *, R *, R *, R *, R *, const I *, R *, R *,
                const R *);
void KER(pgas2d_dual)(const R *, R *, R *, R *, const R *, const R *, const R *,
                      const R *, const R *, const I *, const I *, const I *,
                      const I *, const I *, const I *, const R *, const R *);
void KER(calc_dt)(const I *, const I *, const I *, const I *, const I *, const I *,
                  const I *, const I *, const I *, const I *, const I *, const R *,
                  const R *, const R *, const R *, const R *, const R *, const R *,
                  const R *, const I *, const R *, const R *, const R *, const R *,
                  const R *, const R *, R *, R *);

void KER(expand_terms)(const I *, const I *, const I *, const R *, const I *, const R *, R *,
                       const R *, R *, R *, R *, R *, R *, const R *, const R *, const R *,
                       const R *, const R *, const R *, const I *, R *, const R *);

/* ---------------------------------------------------------------- */
/* grid::ComputePressure, time == Time branch (Grid_ComputePressure.C:96-166),
   EquationOfState 0, EOSType 0, no MHD / cosmic rays. */
int PFX(compute_pressure)(const enzo_ppm_grid *g, const enzo_ppm_globals *gl,
                          R MinimumSupportEnergyCoefficient, R *pressure) {
  I size = 1;
  for (I d = 0; d < g->GridRank; d++) size *= g->GridDimension[d];
  R **bf = g->BaryonField;
  const R OneHalf = 0.5;
  for (I i = 0; i < size; i++) {
    R total_energy = bf[g->TENum][i];
    R density = bf[g->DensNum][i];
    R velocity1 = bf[g->Vel1Num][i];
    R velocity2 = g->GridRank > 1 ? bf[g->Vel2Num][i] : 0.0;
    R velocity3 = g->GridRank > 2 ? bf[g->Vel3Num][i] : 0.0;
    R gas_energy;
    if (gl->DualEnergyFormalism == 0)
      gas_energy = total_energy - OneHalf * (velocity1 * velocity1 + velocity2 * velocity2 +
                                             velocity3 * velocity3);
    else
      gas_energy = bf[g->GENum][i];
    pressure[i] = (gl->Gamma - 1.0) * density * gas_energy;
    if (pressure[i] < gl->tiny_number) pressure[i] = gl->tiny_number;
    R min_pressure = MinimumSupportEnergyCoefficient * (gl->Gamma - 1.0) * density * density;
    if (pressure[i] < min_pressure) pressure[i] = min_pressure;
  }
  return ENZO_PPM_SUCCESS;
}

/* grid::ComputePressure at an interpolated time (Grid_ComputePressure.C:
   56-61 and the general branch :170-240): every field the pressure reads
   is coef * BaryonField + (1 - coef) * OldBaryonField, coef being
   (time - OldTime) / (Time - OldTime).  EquationOfState 0, EOSType 0. */
int PFX(compute_pressure_at)(const enzo_ppm_grid *g, R *const *OldBaryonField,
                             const enzo_ppm_globals *gl, R coef,
                             R MinimumSupportEnergyCoefficient, R *pressure) {
  I size = 1;
  for (I d = 0; d < g->GridRank; d++) size *= g->GridDimension[d];
  R **bf = g->BaryonField;
  R *const *ob = OldBaryonField;
  const R coefold = 1 - coef;
  const R OneHalf = 0.5;
  for (I i = 0; i < size; i++) {
    R total_energy = coef * bf[g->TENum][i] + coefold * ob[g->TENum][i];
    R density = coef * bf[g->DensNum][i] + coefold * ob[g->DensNum][i];
    R velocity1 = coef * bf[g->Vel1Num][i] + coefold * ob[g->Vel1Num][i];
    R velocity2 = 0, velocity3 = 0;
    if (g->GridRank > 1) velocity2 = coef * bf[g->Vel2Num][i] + coefold * ob[g->Vel2Num][i];
    if (g->GridRank > 2) velocity3 = coef * bf[g->Vel3Num][i] + coefold * ob[g->Vel3Num][i];
    R gas_energy;
    if (gl->DualEnergyFormalism == 0) {
      R kineticE = OneHalf * (velocity1 * velocity1 + velocity2 * velocity2 +
                              velocity3 * velocity3);
      gas_energy = total_energy - kineticE;
    } else {
      gas_energy = coef * bf[g->GENum][i] + coefold * ob[g->GENum][i];
    }
    pressure[i] = (gl->Gamma - 1.0) * density * gas_energy;
    if (pressure[i] < gl->tiny_number) pressure[i] = gl->tiny_number;
    R min_pressure = MinimumSupportEnergyCoefficient * (gl->Gamma - 1.0) * density * density;
    if (pressure[i] < min_pressure) pressure[i] = min_pressure;
  }
  return ENZO_PPM_SUCCESS;
}

/* grid::ComovingExpansionTerms (Grid_ComovingExpansionTerms.C:55-130):
   the pressure at PressureTime -- the interpolated form with weight `coef`
   when an old state exists (OldBaryonField non-NULL and coef != 1: the
   reference's time != Time), else the time == Time form -- then
   expand_terms over every cell with Coefficient = dtFixed * dadt / a,
   imethod = HydroMethod (0, PPM direct Euler), no cosmic rays. */
int PFX(comoving_expansion_terms)(const enzo_ppm_grid *g, R *const *OldBaryonField,
                                  const enzo_ppm_globals *gl, R coef, R Coefficient) {
  I size = 1;
  for (I d = 0; d < g->GridRank; d++) size *= g->GridDimension[d];
  R *p = (R *)malloc(sizeof(R) * (size_t)size);
  if (!p) return ENZO_PPM_ERR_ARGUMENT;
  if (OldBaryonField && coef != 1.0)
    PFX(compute_pressure_at)(g, OldBaryonField, gl, coef, 0.0, p);
  else
    PFX(compute_pressure)(g, gl, 0.0, p);
  R **bf = g->BaryonField;
  R *d = bf[g->DensNum];
  R *ge = gl->DualEnergyFormalism ? bf[g->GENum] : d;
  R *v = g->GridRank > 1 ? bf[g->Vel2Num] : bf[g->Vel1Num];
  R *w = g->GridRank > 2 ? bf[g->Vel3Num] : bf[g->Vel1Num];
  I rank = g->GridRank, idual = gl->DualEnergyFormalism, imethod = 0, icr = 0;
  R ecr = 0, ecrold = 0;
  KER(expand_terms)(&rank, &size, &idual, &Coefficient, &imethod, &gl->Gamma, p, d,
                    bf[g->TENum], ge, bf[g->Vel1Num], v, w, d, d, d, d, d, d, &icr, &ecr,
                    &ecrold);
  free(p);
  return ENZO_PPM_SUCCESS;
}

/* calc_dt call of grid::ComputeTimeStep (Grid_ComputeTimeStep.C:151-174,209):
   dtBaryons = CourantSafetyNumber * calc_dt(rank, ..., aye).  CellWidth
   arrays are the grid's (not a-scaled) widths; aye is a(Time) in comoving
   coordinates (:103-106), else 1. */
int PFX(compute_time_step)(const enzo_ppm_grid *g, const enzo_ppm_globals *gl,
                           R *CellWidth[3], R CourantSafetyNumber, R aye, R *dt_out) {
  I size = 1;
  for (I d = 0; d < g->GridRank; d++) size *= g->GridDimension[d];
  R *p = (R *)malloc(sizeof(R) * (size_t)size);
  if (!p) return ENZO_PPM_ERR_ARGUMENT;
  PFX(compute_pressure)(g, gl, 0.0, p);
  I dims[3] = {1, 1, 1}, st[3] = {0, 0, 0}, en[3] = {0, 0, 0};
  for (I d = 0; d < g->GridRank; d++) {
    dims[d] = g->GridDimension[d]; st[d] = g->GridStartIndex[d]; en[d] = g->GridEndIndex[d];
  }
  R zero = 0.0, C2 = 0.0, dt = 0, dtv = 0;
  I ihydro = 0; /* PPM_DirectEuler */
  R unit = 1.0;
  R *dx = CellWidth[0], *dy = g->GridRank > 1 ? CellWidth[1] : &unit,
    *dz = g->GridRank > 2 ? CellWidth[2] : &unit;
  R *bf0 = g->BaryonField[g->Vel1Num];
  R *v2 = g->GridRank > 1 ? g->BaryonField[g->Vel2Num] : bf0;
  R *v3 = g->GridRank > 2 ? g->BaryonField[g->Vel3Num] : bf0;
  ERRCODE = 0;
  KER(calc_dt)(&g->GridRank, &dims[0], &dims[1], &dims[2], &st[0], &en[0], &st[1], &en[1],
               &st[2], &en[2], &ihydro, &C2, dx, dy, dz, &zero, &zero, &zero, &gl->Gamma,
               &gl->PressureFree, &aye, g->BaryonField[g->DensNum], p, bf0, v2, v3, &dt, &dtv);
  free(p);
  *dt_out = dt * CourantSafetyNumber;
  return ERRCODE ? ENZO_PPM_ERR_NOT_SUPPORTED : ENZO_PPM_SUCCESS;
}

/* ---------------------------------------------------------------- */
typedef struct {
  R *d, *e, *u, *v, *w, *p, *gr, *ge, *col;
  R *dls, *drs, *flatten, *pbar, *pls, *prs, *ubar, *uls, *urs, *vls, *vrs;
  R *gels, *gers, *wls, *wrs, *diffcoef, *df, *ef, *uf, *vf, *wf, *gef, *ges;
  R *colf, *colls, *colrs;
  R *block;
} slicebuf;

static int slicebuf_alloc(slicebuf *b, I size, I ncol) {
  I nc = ncol > 0 ? ncol : 1;
  size_t n = (size_t)size * (32 + 4 * (size_t)nc);
  b->block = (R *)calloc(n, sizeof(R));
  if (!b->block) return 0;
  R *q = b->block;
#define TAKE(f, cnt) do { b->f = q; q += (size_t)(cnt); } while (0)
  TAKE(d, size); TAKE(e, size); TAKE(u, size); TAKE(v, size); TAKE(w, size);
  TAKE(p, size); TAKE(gr, size); TAKE(ge, size); TAKE(col, nc * size);
  TAKE(dls, size); TAKE(drs, size); TAKE(flatten, size); TAKE(pbar, size);
  TAKE(pls, size); TAKE(prs, size); TAKE(ubar, size); TAKE(uls, size);
  TAKE(urs, size); TAKE(vls, size); TAKE(vrs, size); TAKE(gels, size);
  TAKE(gers, size); TAKE(wls, size); TAKE(wrs, size); TAKE(diffcoef, size);
  TAKE(df, size); TAKE(ef, size); TAKE(uf, size); TAKE(vf, size); TAKE(wf, size);
  TAKE(gef, size); TAKE(ges, size); TAKE(colf, nc * size); TAKE(colls, nc * size);
  TAKE(colrs, nc * size);
#undef TAKE
  return 1;
}

typedef struct {
  const enzo_ppm_grid *g;
  const enzo_ppm_globals *gl;
  I NumberOfSubgrids;
  enzo_fluxes **SubgridFluxes;
  R **CellWidthTemp;
  const I *GridGlobalStart;
  I GravityOn, NumberOfColours;
  const I *colnum;
  R *pressure;
} sweepctx;

/* grid::[xyz]EulerSweep for one slice; dim = 0,1,2; s = k (x), i (y), j (z). */
static I sweep_slice(const sweepctx *c, I dim, I s, slicebuf *b) {
  const enzo_ppm_grid *g = c->g;
  const enzo_ppm_globals *gl = c->gl;
  const I *GD = g->GridDimension;
  R **BF = g->BaryonField;
  const I rank = g->GridRank;
  const I nxz = g->GridEndIndex[0] - g->GridStartIndex[0] + 1;
  const I nyz = g->GridEndIndex[1] - g->GridStartIndex[1] + 1;
  const I nzz = g->GridEndIndex[2] - g->GridStartIndex[2] + 1;
  const I ncol = c->NumberOfColours;
  const R MinimumPressure = gl->tiny_number;
  /* slice geometry: idim along the sweep, jdim the slice's second axis */
  I idim, jdim, sweepaxis = dim, jaxis;
  /* velocity component loaded into u/v/w slices (Grid_[xyz]EulerSweep.C:81-133) */
  I uax, vax, wax;
  if (dim == 0) { idim = GD[0]; jdim = GD[1]; jaxis = 1; uax = 0; vax = 1; wax = 2; }
  else if (dim == 1) { idim = GD[1]; jdim = GD[2]; jaxis = 2; uax = 1; vax = 2; wax = 0; }
  else { idim = GD[2]; jdim = GD[0]; jaxis = 0; uax = 2; vax = 0; wax = 1; }
  (void)jaxis;
  const I size = idim * jdim;
  const I velnum[3] = {g->Vel1Num, g->Vel2Num, g->Vel3Num};
  R *sl_of_ax[3];
  sl_of_ax[uax] = b->u; sl_of_ax[vax] = b->v; sl_of_ax[wax] = b->w;

  /* index3 of slice element (ii along sweep, jj across), 0-based */
#define IDX3(ii, jj)                                                                 \
  (dim == 0 ? ((s * GD[1] + (jj)) * GD[0] + (ii))                                    \
            : dim == 1 ? (((jj) * GD[1] + (ii)) * GD[0] + s)                         \
                       : (((ii) * GD[1] + s) * GD[0] + (jj)))
  for (I jj = 0; jj < jdim; jj++)
    for (I ii = 0; ii < idim; ii++) {
      I i3 = IDX3(ii, jj), i2 = jj * idim + ii;
      b->d[i2] = BF[g->DensNum][i3];
      b->e[i2] = BF[g->TENum][i3];
      b->p[i2] = c->pressure[i3];
      for (I a = 0; a < 3; a++) sl_of_ax[a][i2] = (a < rank) ? BF[velnum[a]][i3] : 0.0;
      if (c->GravityOn) b->gr[i2] = g->AccelerationField[dim][i3];
      if (gl->DualEnergyFormalism) b->ge[i2] = BF[g->GENum][i3];
      for (I n = 0; n < ncol; n++) b->col[n * size + i2] = BF[c->colnum[n]][i3];
    }

  /* 1-based bounds (Grid_xEulerSweep.C:172-178) */
  I is = g->GridStartIndex[sweepaxis] + 1, ie = g->GridEndIndex[sweepaxis] + 1;
  I js = 1, je = jdim, is_m3 = is - 3, ie_p1 = ie + 1, ie_p3 = ie + 3, k_p1 = s + 1;
  I dim_p1 = dim + 1;
  R **CW = c->CellWidthTemp;
  R dt = g->dtFixed;

  ERRCODE = 0;
  if (gl->PPMDiffusionParameter != 0 || gl->PPMFlatteningParameter != 0) {
    /* calcdiss argument permutation (Grid_[xyz]EulerSweep.C:196-205) */
    const R *v3, *w3, *ddx, *ddy, *ddz;
    const I *pid, *pjd, *pkd, *pn;
    if (dim == 0) {
      v3 = BF[g->Vel2Num]; w3 = BF[g->Vel3Num]; ddx = CW[0]; ddy = CW[1]; ddz = CW[2];
      pid = &GD[0]; pjd = &GD[1]; pkd = &GD[2]; pn = &nzz;
    } else if (dim == 1) {
      v3 = BF[g->Vel3Num]; w3 = BF[g->Vel1Num]; ddx = CW[1]; ddy = CW[2]; ddz = CW[0];
      pid = &GD[1]; pjd = &GD[2]; pkd = &GD[0]; pn = &nxz;
    } else {
      v3 = BF[g->Vel1Num]; w3 = BF[g->Vel2Num]; ddx = CW[2]; ddy = CW[0]; ddz = CW[1];
      pid = &GD[2]; pjd = &GD[0]; pkd = &GD[1]; pn = &nyz;
    }
    KER(calcdiss)(b->d, b->e, b->u, v3, w3, b->p, ddx, ddy, ddz, pid, pjd, pkd, &is, &ie, &js,
                  &je, &k_p1, pn, &dim_p1, &GD[0], &GD[1], &GD[2], &dt, &gl->Gamma,
                  &gl->PPMDiffusionParameter, &gl->PPMFlatteningParameter, b->diffcoef,
                  b->flatten);
  }
  const R *dxs = CW[dim];
  I GravityOn = c->GravityOn, NumberOfColours = ncol;
  KER(inteuler)(b->d, b->p, &GravityOn, b->gr, b->ge, b->u, b->v, b->w, dxs, b->flatten, &idim,
                &jdim, &is, &ie, &js, &je, &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1,
                &gl->DualEnergyFormalismEta2, &gl->PPMSteepeningParameter,
                &gl->PPMFlatteningParameter, &gl->ConservativeReconstruction,
                &gl->PositiveReconstruction, &dt, &gl->Gamma, &gl->PressureFree, b->dls, b->drs,
                b->pls, b->prs, b->gels, b->gers, b->uls, b->urs, b->vls, b->vrs, b->wls,
                b->wrs, &NumberOfColours, b->col, b->colls, b->colrs);
  if (gl->RiemannSolver == ENZO_PPM_RIEMANN_HLLC) {
    /* Grid_xEulerSweep.C:258-269: HLLC fluxes over the whole slice */
    KER(flux_hllc)(b->d, b->e, b->ge, b->u, b->v, b->w, dxs, b->diffcoef, &idim, &jdim, &is, &ie,
                   &js, &je, &dt, &gl->Gamma, &gl->PPMDiffusionParameter,
                   &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1,
                   &gl->RiemannSolverFallback, b->dls, b->drs, b->pls, b->prs, b->uls, b->urs,
                   b->vls, b->vrs, b->wls, b->wrs, b->gels, b->gers, b->df, b->uf, b->vf, b->wf,
                   b->ef, b->gef, b->ges, &NumberOfColours, b->col, b->colls, b->colrs, b->colf);
  } else if (gl->RiemannSolver == ENZO_PPM_RIEMANN_HLL) {
    /* Grid_xEulerSweep.C:245-256: HLL fluxes over the whole slice */
    KER(flux_hll)(b->d, b->e, b->ge, b->u, b->v, b->w, dxs, b->diffcoef, &idim, &jdim, &is, &ie,
                  &js, &je, &dt, &gl->Gamma, &gl->PPMDiffusionParameter,
                  &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1,
                  &gl->RiemannSolverFallback, b->dls, b->drs, b->pls, b->prs, b->uls, b->urs,
                  b->vls, b->vrs, b->wls, b->wrs, b->gels, b->gers, b->df, b->uf, b->vf, b->wf,
                  b->ef, b->gef, b->ges, &NumberOfColours, b->col, b->colls, b->colrs, b->colf);
  } else {
  KER(twoshock)(b->dls, b->drs, b->pls, b->prs, b->uls, b->urs, &idim, &jdim, &is, &ie_p1, &js,
                &je, &dt, &gl->Gamma, &MinimumPressure, &gl->PressureFree, b->pbar, b->ubar,
                &GravityOn, b->gr, &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1);
  KER(flux_twoshock)(b->d, b->e, b->ge, b->u, b->v, b->w, dxs, b->diffcoef, &idim, &jdim, &is,
                     &ie, &js, &je, &dt, &gl->Gamma, &gl->PPMDiffusionParameter,
                     &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1,
                     &gl->RiemannSolverFallback, b->dls, b->drs, b->pls, b->prs, b->gels,
                     b->gers, b->uls, b->urs, b->vls, b->vrs, b->wls, b->wrs, b->pbar, b->ubar,
                     b->df, b->ef, b->uf, b->vf, b->wf, b->gef, b->ges, &NumberOfColours, b->col,
                     b->colls, b->colrs, b->colf);
  }
  KER(euler)(b->d, b->e, b->gr, b->ge, b->u, b->v, b->w, dxs, b->diffcoef, &idim, &jdim, &is,
             &ie, &js, &je, &dt, &gl->Gamma, &gl->PPMDiffusionParameter, &GravityOn,
             &gl->DualEnergyFormalism, &gl->DualEnergyFormalismEta1,
             &gl->DualEnergyFormalismEta2, b->df, b->ef, b->uf, b->vf, b->wf, b->gef, b->ges,
             &NumberOfColours, b->col, b->colf, &gl->SmallRho);
  if (gl->DualEnergyFormalism)
    KER(pgas2d_dual)(b->d, b->e, b->ge, b->p, b->u, b->v, b->w, &gl->DualEnergyFormalismEta1,
                     &gl->DualEnergyFormalismEta2, &idim, &jdim, &is_m3, &ie_p3, &js, &je,
                     &gl->Gamma, &MinimumPressure);

  /* face fluxes of the subgrids (Grid_xEulerSweep.C:313-376 and y/z) */
  I idimax = dim == 0 ? 1 : 0, jdimax = dim == 2 ? 1 : 2;
  for (I n = 0; n < c->NumberOfSubgrids; n++) {
    enzo_fluxes *F = c->SubgridFluxes[n];
    I fistart = F->RightFluxStartGlobalIndex[dim][idimax] - c->GridGlobalStart[idimax];
    I fiend = F->RightFluxEndGlobalIndex[dim][idimax] - c->GridGlobalStart[idimax];
    I fjstart = F->RightFluxStartGlobalIndex[dim][jdimax] - c->GridGlobalStart[jdimax];
    I fjend = F->RightFluxEndGlobalIndex[dim][jdimax] - c->GridGlobalStart[jdimax];
    I nfi = fiend - fistart + 1;
    I lface = F->LeftFluxStartGlobalIndex[dim][dim] - c->GridGlobalStart[dim];
    I rface = F->RightFluxStartGlobalIndex[dim][dim] - c->GridGlobalStart[dim] + 1;
    /* the slice index s runs over the outer axis; the loop runs over the slice row */
    I outer_lo, outer_hi, row_lo, row_hi;
    if (dim == 0) { outer_lo = fjstart; outer_hi = fjend; row_lo = fistart; row_hi = fiend; }
    else if (dim == 1) { outer_lo = fistart; outer_hi = fiend; row_lo = fjstart; row_hi = fjend; }
    else { outer_lo = fjstart; outer_hi = fjend; row_lo = fistart; row_hi = fiend; }
    if (!(s >= outer_lo && s <= outer_hi)) continue;
    for (I r = row_lo; r <= row_hi; r++) {
      I offset = dim == 1 ? (s - fistart) + (r - fjstart) * nfi
                          : (r - fistart) + (s - fjstart) * nfi;
      I lindex = r * GD[dim] + lface, rindex = r * GD[dim] + rface;
      F->LeftFluxes[g->DensNum][dim][offset] = b->df[lindex];
      F->RightFluxes[g->DensNum][dim][offset] = b->df[rindex];
      F->LeftFluxes[g->TENum][dim][offset] = b->ef[lindex];
      F->RightFluxes[g->TENum][dim][offset] = b->ef[rindex];
      /* velocity fluxes with the per-axis guards of each sweep file */
      R *fl_of_ax[3];
      fl_of_ax[uax] = b->uf; fl_of_ax[vax] = b->vf; fl_of_ax[wax] = b->wf;
      const I active[3] = {nxz, nyz, nzz};
      for (I a = 0; a < 3; a++) {
        if (a != dim && active[a] <= 1) continue;
        F->LeftFluxes[velnum[a]][dim][offset] = fl_of_ax[a][lindex];
        F->RightFluxes[velnum[a]][dim][offset] = fl_of_ax[a][rindex];
      }
      if (gl->DualEnergyFormalism) {
        F->LeftFluxes[g->GENum][dim][offset] = b->gef[lindex];
        F->RightFluxes[g->GENum][dim][offset] = b->gef[rindex];
      }
      for (I nc = 0; nc < ncol; nc++) {
        I cl = (r + nc * jdim) * GD[dim] + lface, cr = (r + nc * jdim) * GD[dim] + rface;
        F->LeftFluxes[c->colnum[nc]][dim][offset] = b->colf[cl];
        F->RightFluxes[c->colnum[nc]][dim][offset] = b->colf[cr];
      }
    }
  }

  /* scatter back (Grid_xEulerSweep.C:379-419) */
  for (I jj = 0; jj < jdim; jj++)
    for (I ii = 0; ii < idim; ii++) {
      I i3 = IDX3(ii, jj), i2 = jj * idim + ii;
      BF[g->DensNum][i3] = b->d[i2];
      BF[g->TENum][i3] = b->e[i2];
      for (I a = 0; a < 3; a++)
        if (a < rank) BF[velnum[a]][i3] = sl_of_ax[a][i2];
      if (gl->DualEnergyFormalism) BF[g->GENum][i3] = b->ge[i2];
      for (I n = 0; n < ncol; n++) BF[c->colnum[n]][i3] = b->col[n * size + i2];
    }
#undef IDX3
  return ERRCODE;
}

typedef struct {
  const sweepctx *c;
  I dim, nslices, stride, first;
  I err, fallbacks;
} workerarg;

static void *sweep_worker(void *arg) {
  workerarg *w = (workerarg *)arg;
  const enzo_ppm_grid *g = w->c->g;
  I maxsize = 0;
  for (I a = 0; a < 3; a++)
    for (I b2 = 0; b2 < 3; b2++)
      if (a != b2 && g->GridDimension[a] * g->GridDimension[b2] > maxsize)
        maxsize = g->GridDimension[a] * g->GridDimension[b2];
  slicebuf b;
  if (!slicebuf_alloc(&b, maxsize, w->c->NumberOfColours)) { w->err = -1; return NULL; }
  FALLBACKS = 0;
  for (I s = w->first; s < w->nslices; s += w->stride) {
    I e = sweep_slice(w->c, w->dim, s, &b);
    if (e && !w->err) w->err = e;
  }
  w->fallbacks = FALLBACKS;
  free(b.block);
  return NULL;
}

static I run_sweep(const sweepctx *c, I dim, I nthreads, I *fallbacks) {
  I nslices = c->g->GridDimension[dim == 0 ? 2 : (dim == 1 ? 0 : 1)];
  if (c->gl->PPMDiffusionParameter != 0 || nthreads <= 1) nthreads = 1;
  if (nthreads > nslices) nthreads = nslices;
  workerarg wa[256];
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)64 << 20);
  for (I t = 0; t < nthreads; t++) {
    wa[t].c = c; wa[t].dim = dim; wa[t].nslices = nslices; wa[t].stride = nthreads;
    wa[t].first = t; wa[t].err = 0; wa[t].fallbacks = 0;
    pthread_create(&th[t], &attr, sweep_worker, &wa[t]);
  }
  I err = 0;
  for (I t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    if (wa[t].err && !err) err = wa[t].err;
    *fallbacks += wa[t].fallbacks;
  }
  pthread_attr_destroy(&attr);
  return err;
}

/* grid::SolvePPM_DE (Grid_SolvePPM_DE.C:36-164).  Returns ENZO_PPM_SUCCESS or
   an error code; *fallback_count receives the number of HLL fallback cells. */
int PFX(solve_ppm_de)(const enzo_ppm_grid *g, const enzo_ppm_globals *gl, I CycleNumber,
                      I NumberOfSubgrids, enzo_fluxes **SubgridFluxes, R **CellWidthTemp,
                      const I *GridGlobalStart, I GravityOn, I NumberOfColours,
                      const I *colnum, R MinimumSupportEnergyCoefficient, I nthreads,
                      I *fallback_count) {
  if ((gl->RiemannSolver != ENZO_PPM_RIEMANN_TWOSHOCK && gl->RiemannSolver != ENZO_PPM_RIEMANN_HLL &&
       gl->RiemannSolver != ENZO_PPM_RIEMANN_HLLC) ||
      gl->ReconstructionMethod != ENZO_PPM_RECONSTRUCTION_PPM)
    return ENZO_PPM_ERR_NOT_SUPPORTED;
  if (NumberOfColours > ENZO_PPM_MAX_COLOR) return ENZO_PPM_ERR_ARGUMENT;
  if (gl->EOSType != 0) return ENZO_PPM_ERR_NOT_SUPPORTED; /* :156-157 */
  const I nxz = g->GridEndIndex[0] - g->GridStartIndex[0] + 1;
  const I nyz = g->GridEndIndex[1] - g->GridStartIndex[1] + 1;
  const I nzz = g->GridEndIndex[2] - g->GridStartIndex[2] + 1;
  const I ixyz = CycleNumber % g->GridRank;
  I size = 1;
  for (I d = 0; d < g->GridRank; d++) size *= g->GridDimension[d];
  R *pressure = (R *)calloc((size_t)size, sizeof(R));
  if (!pressure) return ENZO_PPM_ERR_ARGUMENT;
  PFX(compute_pressure)(g, gl, MinimumSupportEnergyCoefficient, pressure);
  sweepctx c = {g, gl, NumberOfSubgrids, SubgridFluxes, CellWidthTemp, GridGlobalStart,
                GravityOn, NumberOfColours, colnum, pressure};
  I fb = 0, err = 0;
  for (I n = ixyz; n < ixyz + g->GridRank && !err; n++) {
    I dir = n % g->GridRank;
    if (dir == 0 && nxz > 1) err = run_sweep(&c, 0, nthreads, &fb);
    if (dir == 1 && nyz > 1) err = run_sweep(&c, 1, nthreads, &fb);
    if (dir == 2 && nzz > 1) err = run_sweep(&c, 2, nthreads, &fb);
  }
  free(pressure);
  if (fallback_count) *fallback_count = fb;
  if (err) return err == 2 || err == 3 || err == 8 ? ENZO_PPM_ERR_NOT_SUPPORTED
                                                   : ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}

/* One directional sweep, grid::[xyz]EulerSweep over every slice of `dim`
   (Grid_SolvePPM_DE.C:88-144 for one direction), with the caller's pressure
   array standing for the stale pressure of the step.  Lets tests compose a
   step from single sweeps in an order they state themselves. */
int PFX(sweep_direction)(const enzo_ppm_grid *g, const enzo_ppm_globals *gl, I dim,
                         I NumberOfSubgrids, enzo_fluxes **SubgridFluxes, R **CellWidthTemp,
                         const I *GridGlobalStart, I GravityOn, I NumberOfColours,
                         const I *colnum, R *pressure, I nthreads, I *fallback_count) {
  if (dim < 0 || dim >= g->GridRank || !pressure) return ENZO_PPM_ERR_ARGUMENT;
  if (NumberOfColours > ENZO_PPM_MAX_COLOR) return ENZO_PPM_ERR_ARGUMENT;
  sweepctx c = {g, gl, NumberOfSubgrids, SubgridFluxes, CellWidthTemp, GridGlobalStart,
                GravityOn, NumberOfColours, colnum, pressure};
  I fb = 0;
  I err = run_sweep(&c, dim, nthreads, &fb);
  if (fallback_count) *fallback_count = fb;
  if (err) return err == 2 || err == 3 || err == 8 ? ENZO_PPM_ERR_NOT_SUPPORTED
                                                   : ENZO_PPM_ERR_NEGATIVE;
  return ENZO_PPM_SUCCESS;
}
