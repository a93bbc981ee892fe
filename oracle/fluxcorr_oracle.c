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
 *     for a proper subgrid (SUBlingGrid FALSE, UpdateFromFinerGrids.C:
 *     155-178) and for a SUBling (SUBlingGrid TRUE, UpdateFromFinerGrids.C:
 *     180-198: a finer grid that is not this grid's child but touches one
 *     of its faces), Cartesian coordinates, HydroMethod PPM_DirectEuler,
 *     RadiativeCooling 0, FluxCorrection 1.
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
 * invariants, the undo rule, the projection's sums, and a SUBling's
 * correction equal to the proper-subgrid correction of one grid holding
 * both neighbours, across a periodic wrap too).
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
 * supported cases).  GridGlobalStart: global index of the grid's first cell,
 * ghosts included (nlongint((CellLeftEdge[i][0] - DomainLeftEdge[i]) /
 * CellWidth[i][0])).  Boundary may be NULL (then CorrectParentBoundaryFlux
 * must be 0).  SUBlingGrid: Initial is this grid's own boundary fluxes and
 * Refined a SUBling's projected fluxes; GlobalDim[dim] is the level's
 * domain size in cells (nlongint((DomainRightEdge - DomainLeftEdge) /
 * CellWidth)) and PeriodicFace[2 dim + side] is 1 where the domain's
 * left/right face along dim is periodic (:147-216).  The reference edits
 * Refined's index ranges for a SUBling in place; here they are edited in a
 * copy.  Refined's flux values are reset where corrections are undone, as
 * in the reference.  Returns ENZO_PPM_ERR_ARGUMENT where the reference
 * fails. */
int orc_correct_for_refined_fluxes(const enzo_ppm_grid *g, const I *GridGlobalStart,
                                   const enzo_fluxes *Initial, enzo_fluxes *Refined,
                                   enzo_fluxes *Boundary, I CorrectParentBoundaryFlux,
                                   I DualEnergyFormalism, R tiny_number, I ncolours,
                                   const I *colnum, const I *colkind, I SUBlingGrid,
                                   const I *GlobalDim, const I *PeriodicFace) {
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
  if (SUBlingGrid && (!GlobalDim || !PeriodicFace)) return ENZO_PPM_ERR_ARGUMENT;
  for (I dim = 0; dim < rank; dim++) {
    if (GD[dim] <= 1) continue;
    /* Refined's index ranges of this direction (edited for a SUBling) */
    I RLS[3], RLE[3], RRS[3];
    for (I i = 0; i < 3; i++) {
      RLS[i] = Refined->LeftFluxStartGlobalIndex[dim][i];
      RLE[i] = Refined->LeftFluxEndGlobalIndex[dim][i];
      RRS[i] = Refined->RightFluxStartGlobalIndex[dim][i];
    }
    I RefinedDim[3] = {0, 0, 0}, RefinedOffset[3] = {0, 0, 0}, InitialOffset[3] = {0, 0, 0};
    I CorrectLeftBaryonField = 1, CorrectRightBaryonField = 1;
    if (!SUBlingGrid) {
      for (I j = 0; j < rank; j++)
        if (Initial->LeftFluxStartGlobalIndex[dim][j] != RLS[j] ||
            Initial->LeftFluxEndGlobalIndex[dim][j] != RLE[j])
          return ENZO_PPM_ERR_ARGUMENT;
    } else {
      /* :147-222: the SUBling's faces moved one cell outwards (wrapped on a
       * periodic domain face); the face that meets this grid's own face is
       * the one corrected */
      const I G = GlobalDim[dim];
      for (I i = 0; i < 3; i++) RefinedDim[i] = RLE[i] - RLS[i] + 1;
      if (RLS[dim] == 0 && PeriodicFace[2 * dim]) {
        RLS[dim] = G - 1;
        RLE[dim] = G - 1;
      } else {
        RLS[dim]--;
        RLE[dim]--;
      }
      if (RRS[dim] == G - 1 && PeriodicFace[2 * dim + 1])
        RRS[dim] = 0;
      else
        RRS[dim]++;
      CorrectRightBaryonField = Initial->RightFluxStartGlobalIndex[dim][dim] == RLS[dim];
      CorrectLeftBaryonField = Initial->LeftFluxStartGlobalIndex[dim][dim] == RRS[dim];
      for (I i = 0; i < 3; i++)
        RefinedOffset[i] = imax(Initial->LeftFluxStartGlobalIndex[dim][i] - RLS[i], 0);
      RLS[dim] = 0;
      RefinedOffset[dim] = 0;
    }
    if (!CorrectLeftBaryonField && !CorrectRightBaryonField) continue;
    I Start[3] = {0, 0, 0}, End[3] = {0, 0, 0}, Dim[3];
    for (I i = 0; i < rank; i++) {
      Start[i] = imax(Initial->LeftFluxStartGlobalIndex[dim][i], RLS[i]) - GridGlobalStart[i];
      End[i] = imin(Initial->LeftFluxEndGlobalIndex[dim][i], RLE[i]) - GridGlobalStart[i];
      if (Start[i] < 0 || End[i] > GD[i]) return ENZO_PPM_ERR_ARGUMENT;
    }
    Start[dim] = imax(Start[dim] - 1, 0);
    End[dim] = Start[dim];
    for (I i = 0; i < 3; i++) Dim[i] = End[i] - Start[i] + 1;
    I Offset = Initial->RightFluxStartGlobalIndex[dim][dim] -
               Initial->LeftFluxStartGlobalIndex[dim][dim] + 2;
    Offset = imin(Offset, GD[dim] - 1);
    if (SUBlingGrid) {
      /* :287-303: the corrected cells are this grid's own first / last
       * active layer, and Initial spans the grid's whole face */
      Offset -= 2;
      Start[dim]++;
      End[dim]++;
      for (I i = 0; i < rank; i++) {
        if (i != dim) {
          Dim[i] = g->GridEndIndex[i] - g->GridStartIndex[i] + 1;
          InitialOffset[i] = imax(RLS[i] - Initial->LeftFluxStartGlobalIndex[dim][i], 0);
        } else {
          Dim[i] = 1;
          InitialOffset[i] = 0;
        }
      }
    }
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
            const I fi = ((k - Start[2] + InitialOffset[2]) * Dim[1] + (j - Start[1] + InitialOffset[1])) *
                             Dim[0] + (i - Start[0] + InitialOffset[0]);
            const I rfi = SUBlingGrid ? ((k - Start[2] + RefinedOffset[2]) * RefinedDim[1] +
                                         (j - Start[1] + RefinedOffset[1])) * RefinedDim[0] +
                                            (i - Start[0] + RefinedOffset[0])
                                      : fi;
            const I gfi = (i - GFS[0]) + (j - GFS[1]) * GFD[0] + (k - GFS[2]) * GFD[1] * GFD[0];
            R *IL = Initial->LeftFluxes[f][dim], *IR = Initial->RightFluxes[f][dim];
            R *RL = Refined->LeftFluxes[f][dim], *RR = Refined->RightFluxes[f][dim];
            if (CorrectLeftBoundaryFlux && CorrectParentBoundaryFlux)
              Boundary->LeftFluxes[f][dim][gfi] = RL[fi];
            if (CorrectLeftBaryonField) {
              if (!SUBlingGrid) BF[f][x] += IL[fi] - RL[fi];
              else BF[f][x] += -(IL[fi] - RR[rfi]);
            }
            if (CorrectRightBoundaryFlux && CorrectParentBoundaryFlux)
              Boundary->RightFluxes[f][dim][gfi] = RR[fi];
            if (CorrectRightBaryonField) {
              if (!SUBlingGrid) BF[f][x + Offset] += -(IR[fi] - RR[fi]);
              else BF[f][x + Offset] += IR[fi] - RL[rfi];
            }
            if (!checked) continue;
            if (CorrectLeftBaryonField && BF[f][x] <= 0)
              for (I u = 0; u < nl; u++) {
                const I ff = list[u];
                R *uIL = Initial->LeftFluxes[ff][dim];
                R *uRL = Refined->LeftFluxes[ff][dim], *uRR = Refined->RightFluxes[ff][dim];
                if (!SUBlingGrid) {
                  if (u <= li) BF[ff][x] -= (uIL[fi] - uRL[fi]);
                  uRL[fi] = uIL[fi];
                } else {
                  if (u <= li) BF[ff][x] -= -(uIL[fi] - uRR[rfi]);
                  uRR[rfi] = uIL[fi];
                }
              }
            if (CorrectRightBaryonField && BF[f][x + Offset] <= 0)
              for (I u = 0; u < nl; u++) {
                const I ff = list[u];
                R *uIR = Initial->RightFluxes[ff][dim];
                R *uRL = Refined->LeftFluxes[ff][dim], *uRR = Refined->RightFluxes[ff][dim];
                if (!SUBlingGrid) {
                  if (u <= li) BF[ff][x + Offset] -= -(uIR[fi] - uRR[fi]);
                  uRR[fi] = uIR[fi];
                } else {
                  /* :638-652 subtract the amount of the field being checked
                   * (f), not of ff, from every field already corrected */
                  if (u <= li) BF[ff][x + Offset] -= (IR[fi] - RL[rfi]);
                  uRL[rfi] = uIR[fi];
                }
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
