/* Prints the size of every struct of include/enzo_ppm.h and the offset and
   size of each member, one "struct member offset size" line each (member
   "-" for the struct itself), as a C caller compiled with or without
   ENZO_PPM_INTEGERS_32 sees them; tests/test_int32_abi.py compares the
   lines with the ctypes mirror (enzo_ppm/abi.py) of the same width.
   Test infrastructure only. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/enzo_ppm.h"

#define S(T) printf("%s - 0 %zu\n", #T, sizeof(T))
#define F(T, m) printf("%s %s %zu %zu\n", #T, #m, offsetof(T, m), sizeof(((T *)0)->m))

int main(void) {
  printf("bits %d 0 %zu\n", ENZO_PPM_INTEGER_BITS, sizeof(enzo_int));
  S(enzo_fluxes);
  F(enzo_fluxes, LeftFluxStartGlobalIndex);
  F(enzo_fluxes, LeftFluxEndGlobalIndex);
  F(enzo_fluxes, RightFluxStartGlobalIndex);
  F(enzo_fluxes, RightFluxEndGlobalIndex);
  F(enzo_fluxes, LeftFluxes);
  F(enzo_fluxes, RightFluxes);
  S(enzo_ppm_globals);
  F(enzo_ppm_globals, Gamma);
  F(enzo_ppm_globals, PPMDiffusionParameter);
  F(enzo_ppm_globals, PPMFlatteningParameter);
  F(enzo_ppm_globals, PPMSteepeningParameter);
  F(enzo_ppm_globals, DualEnergyFormalism);
  F(enzo_ppm_globals, DualEnergyFormalismEta1);
  F(enzo_ppm_globals, DualEnergyFormalismEta2);
  F(enzo_ppm_globals, PressureFree);
  F(enzo_ppm_globals, tiny_number);
  F(enzo_ppm_globals, RiemannSolver);
  F(enzo_ppm_globals, RiemannSolverFallback);
  F(enzo_ppm_globals, ReconstructionMethod);
  F(enzo_ppm_globals, ConservativeReconstruction);
  F(enzo_ppm_globals, PositiveReconstruction);
  F(enzo_ppm_globals, SmallRho);
  F(enzo_ppm_globals, EOSType);
  S(enzo_ppm_grid);
  F(enzo_ppm_grid, GridRank);
  F(enzo_ppm_grid, GridDimension);
  F(enzo_ppm_grid, GridStartIndex);
  F(enzo_ppm_grid, GridEndIndex);
  F(enzo_ppm_grid, BaryonField);
  F(enzo_ppm_grid, AccelerationField);
  F(enzo_ppm_grid, DensNum);
  F(enzo_ppm_grid, GENum);
  F(enzo_ppm_grid, Vel1Num);
  F(enzo_ppm_grid, Vel2Num);
  F(enzo_ppm_grid, Vel3Num);
  F(enzo_ppm_grid, TENum);
  F(enzo_ppm_grid, dtFixed);
  S(hipPPMParameter);
  F(hipPPMParameter, DensNum);
  F(hipPPMParameter, TENum);
  F(hipPPMParameter, Vel1Num);
  F(hipPPMParameter, Vel2Num);
  F(hipPPMParameter, Vel3Num);
  F(hipPPMParameter, GENum);
  F(hipPPMParameter, MaxNumberOfBaryonFields);
  F(hipPPMParameter, GridGlobalStart);
  F(hipPPMParameter, GridStartIndex);
  F(hipPPMParameter, GridEndIndex);
  F(hipPPMParameter, GridDimension);
  F(hipPPMParameter, GridRank);
  F(hipPPMParameter, dx);
  F(hipPPMParameter, CellWidth);
  F(hipPPMParameter, Gamma);
  F(hipPPMParameter, PPMDiffusionParameter);
  F(hipPPMParameter, PPMFlatteningParameter);
  F(hipPPMParameter, PPMSteepeningParameter);
  F(hipPPMParameter, ConservativeReconstruction);
  F(hipPPMParameter, PositiveReconstruction);
  F(hipPPMParameter, GravityOn);
  F(hipPPMParameter, DualEnergyFormalism);
  F(hipPPMParameter, DualEnergyFormalismEta1);
  F(hipPPMParameter, DualEnergyFormalismEta2);
  F(hipPPMParameter, PressureFree);
  F(hipPPMParameter, MinimumPressure);
  F(hipPPMParameter, RiemannSolver);
  F(hipPPMParameter, RiemannSolverFallback);
  F(hipPPMParameter, SmallRho);
  F(hipPPMParameter, NumberOfColours);
  F(hipPPMParameter, colnum);
  return 0;
}
