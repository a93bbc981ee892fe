/*
 * ORACLE -- test infrastructure only.
 *
 * C restatement of the AMR flux correction that consumes the SubgridFluxes
 * a hydro step produces:
 *   - grid::GetProjectedBoundaryFluxes (Grid_GetProjectedBoundaryFluxes.C:
 *     43-125): a subgrid's boundary fluxes summed onto the parent's
 *     resolution, each fine value times 1 / (product of the refinement
 *     factors), in the fine loop order (k, j, i);
 *   - grid::CorrectForRefinedFluxes (Grid_CorrectForRefinedFluxes.C:53-870)
 *     for a proper subgrid (SUBlingGrid == FALSE), Cartesian coordinates,
 *     HydroMethod PPM_DirectEuler, RadiativeCooling 0, FluxCorrection 1.
 *
 * Fields: density, total energy, gas energy (with dual energy) and the three
 * velocities play their roles; each colour field has a kind:
 *   0 species-like (FieldType ElectronDensity..ExtraType1 or MetalSNIa/II):
 *     turned into a fraction of the density before the correction, not
 *     corrected, multiplied by the new density after it;
 *   1 passive: untouched;
 *   2 conserved: multiplied by the density before, divided after, not
 *     corrected (MakeFieldConservative true, not in the corrected list).
 *
 * Only tests/ use this file, as the checker of the library's device version.
 * The reference routine is C++ inside the grid class and cannot be built on
 * its own; what pins this restatement is listed in
 * tests/test_fluxcorr_oracle.py (identity, conservation and fraction
 * invariants, the undo rule, the projection's sums).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "../include/enzo_ppm.h"

typedef int64_t I;
typedef double R;

static inline R fmx(R a, R b) { return a > b ? a : b; }  /* Enzo's max macro */
static inline I imax(I a, I b) { return a > b ? a : b; }
static inline I imin(I a, I b) { return a < b ? a : b; }

/* Grid_GetProjectedBoundaryFluxes.C:57-111 for dims < rank.  The caller
 * provides Projected's arrays (zeroed or not: they are cleared here) and
 * index ranges are set here. */
int orc_project_fluxes(const enzo_fluxes *Fine, I rank, I nfields, const I *Refinement,
                       enzo_fluxes *Projected) {
  for (I dim = 0; dim < rank; dim++)
    for (I i = 0; i < rank; i++) {
      Projected->LeftFluxStartGlobalIndex[dim][i] =
          Fine->LeftFluxStartGlobalIndex[dim][i] / Refinement[i];
      Projected->LeftFluxEndGlobalIndex[dim][i] = Fine->LeftFluxEndGlobalIndex[dim][i] / Refinement[i];
      Projected->RightFluxStartGlobalIndex[dim][i] =
          Fine->RightFluxStartGlobalIndex[dim][i] / Refinement[i];
      Projected->RightFluxEndGlobalIndex[dim][i] =
          Fine->RightFluxEndGlobalIndex[dim][i] / Refinement[i];
    }
  for (I dim = 0; dim < rank; dim++) {
    I Dims[3] = {1, 1, 1}, PD[3];
    for (I i = 0; i < rank; i++)
      Dims[i] = Fine->LeftFluxEndGlobalIndex[dim][i] - Fine->LeftFluxStartGlobalIndex[dim][i] + 1;
    I TotalRefinement = 1;
    for (I i = 0; i < 3; i++) {
      TotalRefinement *= Refinement[i];
      PD[i] = i != dim ? Dims[i] / Refinement[i] : 1;
    }
    const I size = PD[0] * PD[1] * PD[2];
    const R dArea = 1.0 / (R)TotalRefinement;
    for (I f = 0; f < nfields; f++) {
      R *L = Projected->LeftFluxes[f][dim], *Rt = Projected->RightFluxes[f][dim];
      for (I i = 0; i < size; i++) L[i] = Rt[i] = 0.0;
      for (I k = 0; k < Dims[2]; k++) {
        const I k1 = k / Refinement[2];
        for (I j = 0; j < Dims[1]; j++) {
          const I j1 = j / Refinement[1];
          for (I i = 0; i < Dims[0]; i++) {
            const I i1 = i / Refinement[0];
            const I p = i1 + j1 * PD[0] + k1 * PD[0] * PD[1];
            const I q = i + j * Dims[0] + k * Dims[0] * Dims[1];
            L[p] += Fine->LeftFluxes[f][dim][q] * dArea;
            Rt[p] += Fine->RightFluxes[f][dim][q] * dArea;
          }
        }
      }
    }
  }
  return ENZO_PPM_SUCCESS;
}

/* Grid_CorrectForRefinedFluxes.C:53-870 (see the file comment for the
 * supported case).  GridGlobalStart: global index of the grid's first cell,
 * ghosts included (nlongint((CellLeftEdge[i][0] - DomainLeftEdge[i]) /
 * CellWidth[i][0])).  Boundary may be NULL (then CorrectParentBoundaryFlux
 * must be 0).  Returns ENZO_PPM_ERR_ARGUMENT where the reference fails. */
int orc_correct_for_refined_fluxes(const enzo_ppm_grid *g, const I *GridGlobalStart,
                                   const enzo_fluxes *Initial, enzo_fluxes *Refined,
                                   enzo_fluxes *Boundary, I CorrectParentBoundaryFlux,
                                   I DualEnergyFormalism, R tiny_number, I ncolours,
                                   const I *colnum, const I *colkind) {
  R **BF = g->BaryonField;
  const I rank = g->GridRank;
  const I *GD = g->GridDimension;
  /* corrected list: density, total energy, gas energy, then the velocities
   * in field-number order (:301-356) */
  I list[6], nl = 0;
  list[nl++] = g->DensNum;
  list[nl++] = g->TENum;
  if (DualEnergyFormalism) list[nl++] = g->GENum;
  {
    I v[3] = {g->Vel1Num, g->Vel2Num, g->Vel3Num};
    for (I a = 0; a < 3; a++)
      for (I b = a + 1; b < 3; b++)
        if (v[b] < v[a]) { I t = v[a]; v[a] = v[b]; v[b] = t; }
    for (I a = 0; a < 3; a++) list[nl++] = v[a];
  }
  /* fields multiplied by density (MakeFieldConservative): energies,
   * velocities, kind-2 colours; species-like colours: kind 0 */
  I cons[6 + ENZO_PPM_MAX_COLOR], ncons = 0, spec[ENZO_PPM_MAX_COLOR], nspec = 0;
  cons[ncons++] = g->TENum;
  if (DualEnergyFormalism) cons[ncons++] = g->GENum;
  cons[ncons++] = g->Vel1Num;
  cons[ncons++] = g->Vel2Num;
  cons[ncons++] = g->Vel3Num;
  for (I c = 0; c < ncolours; c++) {
    if (colkind[c] == 0) spec[nspec++] = colnum[c];
    else if (colkind[c] == 2) cons[ncons++] = colnum[c];
  }
  const I D = g->DensNum;
  for (I dim = 0; dim < rank; dim++) {
    if (GD[dim] <= 1) continue;
    for (I j = 0; j < rank; j++)
      if (Initial->LeftFluxStartGlobalIndex[dim][j] != Refined->LeftFluxStartGlobalIndex[dim][j] ||
          Initial->LeftFluxEndGlobalIndex[dim][j] != Refined->LeftFluxEndGlobalIndex[dim][j])
        return ENZO_PPM_ERR_ARGUMENT;
    I Start[3] = {0, 0, 0}, End[3] = {0, 0, 0}, Dim[3];
    for (I i = 0; i < rank; i++) {
      Start[i] = imax(Initial->LeftFluxStartGlobalIndex[dim][i],
                      Refined->LeftFluxStartGlobalIndex[dim][i]) - GridGlobalStart[i];
      End[i] = imin(Initial->LeftFluxEndGlobalIndex[dim][i], Refined->LeftFluxEndGlobalIndex[dim][i]) -
               GridGlobalStart[i];
      if (Start[i] < 0 || End[i] > GD[i]) return ENZO_PPM_ERR_ARGUMENT;
    }
    Start[dim] = imax(Start[dim] - 1, 0);
    End[dim] = Start[dim];
    for (I i = 0; i < 3; i++) Dim[i] = End[i] - Start[i] + 1;
    I Offset = Initial->RightFluxStartGlobalIndex[dim][dim] -
               Initial->LeftFluxStartGlobalIndex[dim][dim] + 2;
    Offset = imin(Offset, GD[dim] - 1);
    I CorrectLeftBaryonField = 1, CorrectRightBaryonField = 1;
    I CorrectLeftBoundaryFlux = 0, CorrectRightBoundaryFlux = 0;
    if (Start[dim] == g->GridStartIndex[dim] - 1) {
      CorrectLeftBoundaryFlux = 1;
      CorrectLeftBaryonField = 0;
    }
    if (Start[dim] + Offset == g->GridEndIndex[dim] + 1) {
      CorrectRightBoundaryFlux = 1;
      CorrectRightBaryonField = 0;
    }
    I GFS[3], GFD[3];
    for (I i = 0; i < 3; i++) {
      GFS[i] = g->GridStartIndex[i];
      GFD[i] = g->GridEndIndex[i] - g->GridStartIndex[i] + 1;
    }
    GFS[dim] = Start[dim];
    GFD[dim] = 1;
    for (I i = 0; i < dim; i++) Offset *= GD[i];
    /* faces to conserved form (:358-424) */
    for (I k = Start[2]; k <= End[2]; k++)
      for (I j = Start[1]; j <= End[1]; j++)
        for (I i = Start[0]; i <= End[0]; i++) {
          const I x = (k * GD[1] + j) * GD[0] + i;
          for (I c = 0; c < ncons; c++) {
            BF[cons[c]][x] *= BF[D][x];
            BF[cons[c]][x + Offset] *= BF[D][x + Offset];
          }
          for (I c = 0; c < nspec; c++) {
            BF[spec[c]][x] /= BF[D][x];
            BF[spec[c]][x + Offset] /= BF[D][x + Offset];
          }
        }
    /* correction, list outer, cells inner (:447-671) */
    for (I li = 0; li < nl; li++) {
      const I f = list[li];
      const int checked = li < (DualEnergyFormalism ? 3 : 2);  /* density or energy */
      for (I k = Start[2]; k <= End[2]; k++)
        for (I j = Start[1]; j <= End[1]; j++)
          for (I i = Start[0]; i <= End[0]; i++) {
            const I x = (k * GD[1] + j) * GD[0] + i;
            const I fi = ((k - Start[2]) * Dim[1] + (j - Start[1])) * Dim[0] + (i - Start[0]);
            const I gfi = (i - GFS[0]) + (j - GFS[1]) * GFD[0] + (k - GFS[2]) * GFD[1] * GFD[0];
            if (CorrectLeftBoundaryFlux && CorrectParentBoundaryFlux)
              Boundary->LeftFluxes[f][dim][gfi] = Refined->LeftFluxes[f][dim][fi];
            if (CorrectLeftBaryonField)
              BF[f][x] += Initial->LeftFluxes[f][dim][fi] - Refined->LeftFluxes[f][dim][fi];
            if (CorrectRightBoundaryFlux && CorrectParentBoundaryFlux)
              Boundary->RightFluxes[f][dim][gfi] = Refined->RightFluxes[f][dim][fi];
            if (CorrectRightBaryonField)
              BF[f][x + Offset] +=
                  -(Initial->RightFluxes[f][dim][fi] - Refined->RightFluxes[f][dim][fi]);
            if (!checked) continue;
            if (CorrectLeftBaryonField && BF[f][x] <= 0)
              for (I u = 0; u < nl; u++) {
                const I ff = list[u];
                if (u <= li)
                  BF[ff][x] -= (Initial->LeftFluxes[ff][dim][fi] - Refined->LeftFluxes[ff][dim][fi]);
                Refined->LeftFluxes[ff][dim][fi] = Initial->LeftFluxes[ff][dim][fi];
              }
            if (CorrectRightBaryonField && BF[f][x + Offset] <= 0)
              for (I u = 0; u < nl; u++) {
                const I ff = list[u];
                if (u <= li)
                  BF[ff][x + Offset] -=
                      -(Initial->RightFluxes[ff][dim][fi] - Refined->RightFluxes[ff][dim][fi]);
                Refined->RightFluxes[ff][dim][fi] = Initial->RightFluxes[ff][dim][fi];
              }
          }
    }
    /* back to the original quantities (:779-797) */
    for (I k = Start[2]; k <= End[2]; k++)
      for (I j = Start[1]; j <= End[1]; j++)
        for (I i = Start[0]; i <= End[0]; i++) {
          const I x = (k * GD[1] + j) * GD[0] + i;
          for (I c = 0; c < ncons; c++) {
            BF[cons[c]][x] /= BF[D][x];
            BF[cons[c]][x + Offset] /= BF[D][x + Offset];
          }
        }
    /* dual-energy consistency of the face cells (:801-845) */
    if (DualEnergyFormalism)
      for (I k = Start[2]; k <= End[2]; k++)
        for (I j = Start[1]; j <= End[1]; j++)
          for (I i = Start[0]; i <= End[0]; i++) {
            const I x1 = (k * GD[1] + j) * GD[0] + i, x2 = x1 + Offset;
            R *ge = BF[g->GENum], *te = BF[g->TENum];
            R *v1 = BF[g->Vel1Num], *v2 = BF[g->Vel2Num], *v3 = BF[g->Vel3Num];
            ge[x1] = fmx(ge[x1], tiny_number);
            ge[x2] = fmx(ge[x2], tiny_number);
            te[x1] = ge[x1] + 0.5 * v1[x1] * v1[x1];
            te[x2] = ge[x2] + 0.5 * v1[x2] * v1[x2];
            if (rank > 1) {
              te[x1] += 0.5 * v2[x1] * v2[x1];
              te[x2] += 0.5 * v2[x2] * v2[x2];
            }
            if (rank > 2) {
              te[x1] += 0.5 * v3[x1] * v3[x1];
              te[x2] += 0.5 * v3[x2] * v3[x2];
            }
          }
    /* species back from fractions (:850-867) */
    for (I k = Start[2]; k <= End[2]; k++)
      for (I j = Start[1]; j <= End[1]; j++)
        for (I i = Start[0]; i <= End[0]; i++) {
          const I x = (k * GD[1] + j) * GD[0] + i;
          for (I c = 0; c < nspec; c++) {
            BF[spec[c]][x] *= BF[D][x];
            BF[spec[c]][x + Offset] *= BF[D][x + Offset];
          }
        }
  }
  return ENZO_PPM_SUCCESS;
}
