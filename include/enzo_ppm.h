/*
 * enzo_ppm.h -- C ABI of the MI355X PPM "direct Eulerian" hydro step.
 *
 * Drop-in for the GPU switch inside grid::SolvePPM_DE
 * (src/enzo/Grid_SolvePPM_DE.C:63-82,99-102,118-121,137-141,148-152), which in
 * the reference calls the cuPPM* lifecycle declared in src/enzo/cuPPM.h:53-102.
 * This header exports the same lifecycle for the precision-64 build
 * (`float` = double: src/enzo/macros_and_parameters.h:257) with either
 * integer width (`int` = int64_t by default, int32_t with
 * ENZO_PPM_INTEGERS_32, see enzo_int below), with these additions:
 *   - every call returns a status (ENZO_PPM_SUCCESS, or a negative error
 *     code; the reference calls exit(1) or aborts),
 *   - an explicit hipStream_t,
 *   - CycleNumber drives the x/y/z sweep rotation (Grid_SolvePPM_DE.C:53,86),
 *   - per-cell CellWidthTemp arrays (the reference cuPPM takes one scalar dx
 *     per axis; the Fortran takes dx(idim) arrays),
 *   - SmallRho (the density floor euler.F applies, Grid_xEulerSweep.C:289-296),
 *   - the subgrid face list is registered before the sweeps, so that face
 *     fluxes are captured inside the sweep kernel instead of from full flux
 *     arrays (cuPPM.cu:431-497),
 *   - a one-call entry with grid::SolvePPM_DE's argument list
 *     (Grid.h:1867-1871), and device-resident entries for multi-step runs.
 * No symbol takes a torch, HIP-runtime-owned or C++ type other than
 * hipStream_t (passed as void*).
 */
#ifndef ENZO_PPM_H
#define ENZO_PPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* `int` follows the reference's CONFIG_INTEGERS: 64 by default (LARGE_INTS,
   macros_and_parameters.h:216-217), 32 when the caller defines
   ENZO_PPM_INTEGERS_32 (SMALL_INTS, :201-213; the test-suite build,
   Make.settings.test-suite:2).  Each width is its own library:
   libenzo_ppm_hip.so (64) and libenzo_ppm_hip_i32.so (32), with the same
   entry points; hipPPMVersion() names the width.  long_int is 64 bits in
   both (macros_and_parameters.h:151). */
#ifdef ENZO_PPM_INTEGERS_32
typedef int32_t enzo_int;
#define ENZO_PPM_INTEGER_BITS 32
#else
typedef int64_t enzo_int;
#define ENZO_PPM_INTEGER_BITS 64
#endif
typedef double enzo_float;     /* `float` with precision-64 (macros_and_parameters.h:257) */
typedef int64_t enzo_long_int; /* long_int (macros_and_parameters.h:151) */

#define ENZO_PPM_MAX_DIMENSION 3                 /* MAX_DIMENSION (macros_and_parameters.h:113) */
#ifndef ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS
#define ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS 30  /* PARAMETER_MAX_BARYONS (Make.config.settings:50) */
#endif
#define ENZO_PPM_MAX_COLOR 45                    /* MAX_COLOR (fortran.def:2) */
#define ENZO_PPM_MAX_ANY_SINGLE_DIRECTION 1031   /* fortran.def:1 */

/* status codes: SUCCESS mirrors macros_and_parameters.h:348-349 */
#define ENZO_PPM_SUCCESS 1
#define ENZO_PPM_ERR_ARGUMENT (-1)      /* bad sizes, indices or unsupported option */
#define ENZO_PPM_ERR_HIP (-2)           /* a HIP runtime call failed */
#define ENZO_PPM_ERR_NEGATIVE (-3)      /* negative density/energy with fallback off
                                           (flux_twoshock.F:438-442) or after HLL
                                           (flux_hll.F:336-384) */
#define ENZO_PPM_ERR_NOT_SUPPORTED (-4) /* option the reference rejects on this path */
#define ENZO_PPM_ERR_NO_DEVICE (-5)

/* Riemann solver and reconstruction enums (typedefs.h:278-280) used on this
   path; TwoShock is PPM_DirectEuler's default (ReadParameterFile.C:1574-1578) */
#define ENZO_PPM_RIEMANN_HLL 1
#define ENZO_PPM_RIEMANN_HLLC 4
#define ENZO_PPM_RIEMANN_TWOSHOCK 5
#define ENZO_PPM_RECONSTRUCTION_PPM 1

/* struct fluxes (src/enzo/Fluxes.h:29-37), identical layout */
typedef struct enzo_fluxes {
  enzo_long_int LeftFluxStartGlobalIndex[ENZO_PPM_MAX_DIMENSION][ENZO_PPM_MAX_DIMENSION];
  enzo_long_int LeftFluxEndGlobalIndex[ENZO_PPM_MAX_DIMENSION][ENZO_PPM_MAX_DIMENSION];
  enzo_long_int RightFluxStartGlobalIndex[ENZO_PPM_MAX_DIMENSION][ENZO_PPM_MAX_DIMENSION];
  enzo_long_int RightFluxEndGlobalIndex[ENZO_PPM_MAX_DIMENSION][ENZO_PPM_MAX_DIMENSION];
  enzo_float *LeftFluxes[ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS][ENZO_PPM_MAX_DIMENSION];
  enzo_float *RightFluxes[ENZO_PPM_MAX_NUMBER_OF_BARYON_FIELDS][ENZO_PPM_MAX_DIMENSION];
} enzo_fluxes;

/* Hydro globals read on the path (global_data.h; defaults in
   SetDefaultGlobalValues.C:57,181-195,424-426,803-813). */
typedef struct enzo_ppm_globals {
  enzo_float Gamma;
  enzo_int PPMDiffusionParameter;
  enzo_int PPMFlatteningParameter;
  enzo_int PPMSteepeningParameter;
  enzo_int DualEnergyFormalism;
  enzo_float DualEnergyFormalismEta1;
  enzo_float DualEnergyFormalismEta2;
  enzo_int PressureFree;
  enzo_float tiny_number; /* MinimumPressure in the sweeps */
  enzo_int RiemannSolver;
  enzo_int RiemannSolverFallback;
  enzo_int ReconstructionMethod;
  enzo_int ConservativeReconstruction;
  enzo_int PositiveReconstruction;
  enzo_float SmallRho;
  enzo_int EOSType;
} enzo_ppm_globals;

/* Grid state grid::SolvePPM_DE reads (Grid.h members). */
typedef struct enzo_ppm_grid {
  enzo_int GridRank;
  enzo_int GridDimension[ENZO_PPM_MAX_DIMENSION];
  enzo_int GridStartIndex[ENZO_PPM_MAX_DIMENSION]; /* 0-based, inclusive */
  enzo_int GridEndIndex[ENZO_PPM_MAX_DIMENSION];
  enzo_float **BaryonField;       /* MAX_NUMBER_OF_BARYON_FIELDS pointers, x fastest */
  enzo_float **AccelerationField; /* [3], or NULL without gravity */
  enzo_int DensNum, GENum, Vel1Num, Vel2Num, Vel3Num, TENum;
  enzo_float dtFixed;
} enzo_ppm_grid;

/* cuPPMParameter (cuPPM.h:24-50), FP64/int64, plus the additions above. */
typedef struct hipPPMParameter {
  enzo_int DensNum, TENum, Vel1Num, Vel2Num, Vel3Num, GENum;
  enzo_int MaxNumberOfBaryonFields;
  enzo_long_int GridGlobalStart[3];
  enzo_int GridStartIndex[3];
  enzo_int GridEndIndex[3];
  enzo_int GridDimension[3];
  enzo_int GridRank;
  enzo_float dx[3];                /* uniform widths (cuPPM.cu semantics) */
  const enzo_float *CellWidth[3];  /* optional per-cell widths; NULL = uniform dx */
  enzo_float Gamma;
  enzo_int PPMDiffusionParameter;
  enzo_int PPMFlatteningParameter;
  enzo_int PPMSteepeningParameter;
  enzo_int ConservativeReconstruction;
  enzo_int PositiveReconstruction;
  enzo_int GravityOn;
  enzo_int DualEnergyFormalism;
  enzo_float DualEnergyFormalismEta1;
  enzo_float DualEnergyFormalismEta2;
  enzo_int PressureFree;
  enzo_float MinimumPressure;
  enzo_int RiemannSolver;
  enzo_int RiemannSolverFallback;
  enzo_float SmallRho;
  enzo_int NumberOfColours;
  enzo_int colnum[ENZO_PPM_MAX_COLOR];
} hipPPMParameter;

/* Opaque device state (cuPPMData, cuPPM.h:9-22).  Buffers stay resident
   between calls and are reused while the grid shape is unchanged. */
typedef struct hipPPMData hipPPMData;

/* ---- lifecycle mirroring cuPPM.h:53-102 ---------------------------------- */

/* cuPPMInitParameter (cuPPM.h:53-79); adds GridRank and SmallRho.
   RiemannSolver takes the reference's values (typedefs.h:279-280):
   TwoShock (5, the default of PPM_DirectEuler) with the per-cell HLL
   fallback when RiemannSolverFallback is set, HLL (1): flux_hll over whole
   pencils (Grid_xEulerSweep.C:245-256), its negative-density check giving
   ENZO_PPM_ERR_NEGATIVE, or HLLC (4): flux_hllc (:258-269) with its own
   per-cell flux_hll fallback; with HLL or HLLC each pencil is computed by
   one workgroup that holds it whole, so pencils of up to 1025 cells.  The
   others: ENZO_PPM_ERR_NOT_SUPPORTED.  The batched level
   (hipPPMLevelCreate) takes the same three; with HLL or HLLC every pencil
   group of the level goes to the one-pencil workgroups. */
int hipPPMInitParameter(hipPPMParameter *Para, enzo_int DensNum, enzo_int TENum,
                        enzo_int Vel1Num, enzo_int Vel2Num, enzo_int Vel3Num,
                        enzo_int GENum, enzo_int MaxNumberOfBaryonFields,
                        const enzo_long_int GridGlobalStart[],
                        const enzo_int GridStartIndex[],
                        const enzo_int GridEndIndex[],
                        const enzo_int GridDimension[], enzo_float dxx,
                        enzo_float dxy, enzo_float dxz, enzo_float Gamma,
                        enzo_int PPMDiffusionParameter,
                        enzo_int PPMFlatteningParameter,
                        enzo_int PPMSteepeningParameter,
                        enzo_int ConservativeReconstruction,
                        enzo_int PositiveReconstruction, enzo_int GravityOn,
                        enzo_int DualEnergyFormalism,
                        enzo_float DualEnergyFormalismEta1,
                        enzo_float DualEnergyFormalismEta2,
                        enzo_int PressureFree, enzo_float MinimumPressure,
                        enzo_int RiemannSolver, enzo_int RiemannSolverFallback,
                        enzo_int NumberOfColours, const enzo_int *colnum,
                        enzo_int GridRank, enzo_float SmallRho);

/* Per-cell widths CellWidthTemp[dim][0..GridDimension[dim]-1]
   (Grid_SolveHydroEquations.C:401-409).  The array is copied at InitData. */
int hipPPMSetCellWidth(hipPPMParameter *Para, enzo_int dim, const enzo_float *CellWidthTemp);

/* cuPPMInitData (cuPPM.h:81-83).  *Data may hold a previous handle, which is
   reused when the shape fits.  stream: hipStream_t or NULL (default). */
int hipPPMInitData(hipPPMData **Data, const hipPPMParameter *Para, void *stream);

/* cuPPMSetBaryon (cuPPM.h:88-92): host -> device.  Pressure may be NULL, in
   which case the stale pressure of Grid_SolvePPM_DE.C:58-61 is computed on the
   device (Grid_ComputePressure.C:96-166) with MinimumSupportEnergyCoefficient
   0.  AccelerationField is read only when GravityOn. */
int hipPPMSetBaryon(hipPPMData *Data, const hipPPMParameter *Para,
                    enzo_float *BaryonField[], enzo_float *Pressure,
                    enzo_float *AccelerationField[]);

/* Face list of the subgrids whose fluxes the sweeps capture
   (Grid_xEulerSweep.C:313-376).  Must precede hipPPMSweep for the fluxes to
   be recorded; NumberOfSubgrids = 0 disables capture. */
int hipPPMSetSubgridFluxes(hipPPMData *Data, const hipPPMParameter *Para,
                           enzo_fluxes *SubgridFluxes[], enzo_int NumberOfSubgrids,
                           const enzo_long_int GridGlobalStart[]);

/* cuPPMSweep (cuPPM.h:98-99): one directional sweep, dir 0/1/2 = x/y/z. */
int hipPPMSweep(hipPPMData *Data, const hipPPMParameter *Para, enzo_float dtFixed,
                enzo_int dir);

/* cuPPMSaveSubgridFluxes (cuPPM.h:101-103): copy the fluxes captured in the
   last sweep along dir into the SubgridFluxes arrays registered above. */
int hipPPMSaveSubgridFluxes(enzo_fluxes *SubgridFluxes[], enzo_int NumberOfSubgrids,
                            const enzo_long_int GridGlobalStart[], hipPPMData *Data,
                            const hipPPMParameter *Para, enzo_int dir);

/* cuPPMGetBaryon (cuPPM.h:94-96): device -> host. */
int hipPPMGetBaryon(hipPPMData *Data, const hipPPMParameter *Para,
                    enzo_float *BaryonField[]);

/* cuPPMDestroy (cuPPM.h:85-86). */
int hipPPMDestroy(hipPPMData *Data);

/* Number of cells that took the per-cell HLL fallback
   (flux_twoshock.F:401-456) since the last SetBaryon; -1 on error. */
enzo_long_int hipPPMFallbackCount(hipPPMData *Data);

/* ---- one call with grid::SolvePPM_DE's argument list --------------------- */

/* grid::SolvePPM_DE (Grid_SolvePPM_DE.C:36-164, Grid.h:1867-1871): the
   pressure once, the cycle-rotated sweeps, in-place update of BaryonField and
   the SubgridFluxes fill.  Data may be NULL (a temporary handle is used). */
int hipPPMSolvePPM_DE(const enzo_ppm_grid *grid, const enzo_ppm_globals *globals,
                      enzo_int CycleNumber, enzo_int NumberOfSubgrids,
                      enzo_fluxes *SubgridFluxes[], enzo_float *CellWidthTemp[],
                      const enzo_long_int GridGlobalStart[], enzo_int GravityOn,
                      enzo_int NumberOfColours, const enzo_int colnum[],
                      enzo_float MinimumSupportEnergyCoefficient,
                      hipPPMData **Data, void *stream);

/* ---- the entry point that allocates nothing ------------------------------
   No counterpart in the reference (its GPU path allocates device buffers on
   every call, cuPPM.cu:111-186).  For callers that cannot allocate device
   memory -- tight serving loops, harnesses that forbid hipMalloc: the
   hipPPMSolvePPM_DE step (pressure once, the cycle-rotated sweeps, the
   per-cell fallback) on fields that are already on the current device,
   BaryonField[] and AccelerationField[] holding DEVICE pointers (updated in
   place; CellWidthTemp, colnum and the structs stay on the host).  No face
   fluxes are captured (use hipPPMSolvePPM_DE for subgrid fluxes).  Its
   device buffers come from one static array of
   ENZO_PPM_NOALLOC_SCRATCH_MIB MiB per GPU, held until the process ends:
   a grid needs about 8 * cells * (2 * (6 + NumberOfColours) + 1) bytes of it
   (e.g. up to 121^3 cells, ghosts included, with 3 colours) and a larger one
   is refused with ENZO_PPM_ERR_ARGUMENT.  Every call on a GPU uses that
   GPU's array, so calls that overlap on one GPU (other threads or streams)
   race; calls in turn do not.  Calls no hipMalloc/hipFree; waits for its
   stream before returning.  FallbackCells (may be NULL) receives the cells
   that took the per-cell HLL fallback.  Results are bitwise those of
   hipPPMSolvePPM_DE with NumberOfSubgrids = 0. */
#define ENZO_PPM_NOALLOC_SCRATCH_MIB 256
int hipPPMSolvePPM_DE_NoAlloc(const enzo_ppm_grid *grid, const enzo_ppm_globals *globals,
                              enzo_int CycleNumber, enzo_float *CellWidthTemp[],
                              enzo_int GravityOn, enzo_int NumberOfColours,
                              const enzo_int colnum[],
                              enzo_float MinimumSupportEnergyCoefficient, void *stream,
                              enzo_long_int *FallbackCells);

/* ---- device-resident multi-step use --------------------------------------- */

/* Device pointer of BaryonField[field] (current buffer) or, with field = -1,
   of the pressure array; NULL on error.  Valid until the next sweep. */
enzo_float *hipPPMDeviceField(hipPPMData *Data, enzo_int field);

/* Fused pressure + Courant pass over the device-resident state:
   p = ComputePressure (Grid_ComputePressure.C:96-166) into the pressure
   array, and dt = CourantSafetyNumber * calc_dt (calc_dt.F:129-177,
   Grid_ComputeTimeStep.C:162-209) left on the device for hipPPMStep.
   dt_host, when non-NULL, receives dt (this synchronises the stream). */
int hipPPMPressureAndTimeStep(hipPPMData *Data, const hipPPMParameter *Para,
                              enzo_float CourantSafetyNumber,
                              enzo_float MinimumSupportEnergyCoefficient,
                              enzo_float *dt_host);

/* One full step on the device-resident state: the cycle-rotated x/y/z sweeps
   with the dt left by hipPPMPressureAndTimeStep (dtFixed <= 0) or dtFixed.
   No host synchronisation. */
int hipPPMStep(hipPPMData *Data, const hipPPMParameter *Para, enzo_int CycleNumber,
               enzo_float dtFixed);

/* Periodic ghost-zone fill of every baryon field on the device: what
   SetBoundaryConditions gives a grid that spans a periodic domain, through
   CopyZonesFromGrid from its own periodic images (CheckForOverlap's
   EdgeOffsets; ExternalBoundary::SetExternalBoundary leaves periodic faces
   alone in the reference build, which does not define USE_PERIODIC).
   ENZO_PPM_ERR_ARGUMENT when the ghost zones of an axis differ in width or
   its active extent is below the ghost count (ghost cells more than one
   period from every active cell). */
int hipPPMPeriodicBoundary(hipPPMData *Data, const hipPPMParameter *Para);

/* grid::SetExternalBoundaryValues (Grid_SetExternalBoundaryValues.C:91-117
   -> ExternalBoundary::SetExternalBoundary,
   ExternalBoundary_SetExternalBoundaryIO.C:50-1010, built without
   OOC_BOUNDARY) on the device state, with one boundary type per domain face
   (as ExternalBoundary::InitializeExternalBoundaryFace sets them):
   BoundaryType[2 dim + face] (face 0 left, 1 right) in the reference's
   numbering -- reflecting 0 (mirror of the active cells, the normal
   velocity Velocity1/2/3 negated), outflow 1 (the edge cell), inflow 2
   (InflowValue[field * 6 + 2 dim + face], field = BaryonField number,
   NumberOfBaryonFields rows; NULL without inflow faces), periodic 3 (left
   alone: see hipPPMPeriodicBoundary), shearing 4 (the cell one active extent
   inwards).  GridOffset[dim] = nint((GridLeftEdge - DomainLeftEdge) /
   CellWidth) and BoundaryDimension[dim] (the domain's cells, ghosts
   included) decide which faces lie on the domain's edge: the left one when
   GridOffset is 0, the right one when GridOffset + GridDimension equals
   BoundaryDimension.  Faces in the reference's order (x left, x right, y
   left, ..., z right), each over the grid's whole transverse extent, so
   corners take the later axis' values.  Every field held on the device
   (Vel3 is not held in 2-D).  Other types (hydrostatic, ...), or a shearing
   face whose grid has fewer active cells than ghost layers:
   ENZO_PPM_ERR_NOT_SUPPORTED, with nothing written. */
int hipPPMExternalBoundary(hipPPMData *Data, const hipPPMParameter *Para,
                           const enzo_int *BoundaryType, const enzo_int *GridOffset,
                           const enzo_int *BoundaryDimension, enzo_int NumberOfBaryonFields,
                           const enzo_float *InflowValue);

/* ---- AMR flux correction ------------------------------------------------
   grid::CorrectForRefinedFluxes (Grid_CorrectForRefinedFluxes.C:53-870) on
   the device-resident state, Cartesian coordinates, PPM direct Euler,
   RadiativeCooling 0 and FluxCorrection 1, in 3-D, for the two calls of
   UpdateFromFinerGrids.C:
   - a proper subgrid (SUBlingGrid 0, :155-178): InitialFluxes is the
     subgrid's SubgridFluxes entry of this grid's hydro step, RefinedFluxes
     the subgrid's boundary fluxes projected to this grid
     (grid::GetProjectedBoundaryFluxes); the cells just outside the
     subgrid's faces take the difference; on faces that lie on this grid's
     own boundary the cells are left alone and, with
     CorrectParentBoundaryFlux, BoundaryFluxesThisTimeStep takes the
     refined fluxes;
   - a SUBling (SUBlingGrid 1, :180-198): a finer grid that is not this
     grid's child but touches one of its faces; InitialFluxes is this
     grid's own boundary fluxes, RefinedFluxes the SUBling's fluxes
     projected to this level; this grid's first / last active layer along
     the touching face takes the difference.  GlobalDim[3] is the level's
     domain size in cells and PeriodicFace[6] flags (left, right per axis)
     the periodic domain faces, across which the faces also meet; both may
     be NULL for a proper subgrid.  The reference's undo on a SUBling's
     right-hand face subtracts the amount of the field found non-positive
     from every field corrected before it; this does the same.
   Corrections are made in conserved form, with the reference's undo when
   density or an energy would not stay positive.  ColourKind
   [NumberOfColours]: 0 species-like (FieldType ElectronDensity..ExtraType1,
   MetalSNIa/II: kept as fractions of the density), 1 passive (untouched),
   2 conserved (multiplied by the density around the correction, not
   corrected).  RefinedFluxes is read only (the reference edits its index
   ranges for a SUBling and resets entries it undoes, then frees it).
   ENZO_PPM_ERR_ARGUMENT where the reference fails or a flux index would
   leave the arrays. */
/* grid::GetProjectedBoundaryFluxes (Grid_GetProjectedBoundaryFluxes.C:
   43-111): a subgrid's boundary fluxes (Fine, in its index space) summed
   onto its parent's resolution -- each fine value times 1 / (product of
   the refinement factors), in the reference's summation order -- into
   Projected, whose arrays the caller sizes for the coarse faces; its index
   ranges are set here.  Face extents must be multiples of the factors. */
int hipPPMProjectBoundaryFluxes(const enzo_fluxes *Fine, enzo_int GridRank,
                                enzo_int NumberOfBaryonFields, const enzo_int *RefinementFactors,
                                enzo_fluxes *Projected);

int hipPPMCorrectForRefinedFluxes(hipPPMData *Data, const hipPPMParameter *Para,
                                  const enzo_fluxes *InitialFluxes,
                                  const enzo_fluxes *RefinedFluxes,
                                  enzo_fluxes *BoundaryFluxesThisTimeStep, enzo_int SUBlingGrid,
                                  const enzo_int *GlobalDim, const enzo_int *PeriodicFace,
                                  enzo_int CorrectParentBoundaryFlux, const enzo_int *ColourKind);

/* ---- comoving coordinates ----------------------------------------------
   The reference recasts the solver in comoving form by scaling the cell
   widths it hands to SolvePPM_DE by a(t + dt/2)
   (Grid_SolveHydroEquations.C:386-408), takes the Courant step with a(t)
   (Grid_ComputeTimeStep.C:103-106 -> calc_dt's aye) and adds the expansion
   terms after the step (Grid_ComovingExpansionTerms.C:55-130 ->
   expand_terms.F, ENERGY_METHOD3 and VELOCITY_METHOD3).  On the
   device-resident state:
   hipPPMSetComoving: switch comoving handling on or off.  While on, every
     hipPPMStep first keeps a copy of the state (OldBaryonField), the Courant
     pass uses the unscaled widths times a_courant, the sweeps the widths
     times a_sweep.
   hipPPMSetExpansionFactors: a_courant = a(Time) for hipPPMPressureAndTimeStep,
     a_sweep = a(Time + dt/2) for hipPPMStep (coefficient tables rebuilt when
     it changes).
   hipPPMComovingExpansionTerms: after hipPPMStep, the expansion terms over
     every cell with Coefficient = dt (da/dt) / a at the step's mid-time, the
     pressure taken at interpolation weight `coef` = (PressureTime - OldTime)
     / (Time - OldTime) between the kept copy and the current state (coef = 1:
     the current state, the reference's time == Time branch). */
int hipPPMSetComoving(hipPPMData *Data, const hipPPMParameter *Para, enzo_int enable);
int hipPPMSetExpansionFactors(hipPPMData *Data, const hipPPMParameter *Para, enzo_float a_courant,
                              enzo_float a_sweep);
int hipPPMComovingExpansionTerms(hipPPMData *Data, const hipPPMParameter *Para, enzo_float coef,
                                 enzo_float Coefficient);

/* Timing of the fused sweep kernel: with profiling enabled every sweep
   kernel launch is bracketed by HIP events on the handle's stream;
   hipPPMProfileRead synchronises, returns the summed kernel time (ms) and the
   number of launches since the last read, and resets the record. */
int hipPPMProfile(hipPPMData *Data, enzo_int enable);
int hipPPMProfileRead(hipPPMData *Data, enzo_float *total_ms, enzo_long_int *launches);

/* Wait for all work queued on the handle's stream. */
int hipPPMSynchronize(hipPPMData *Data);

/* ---- learned two-shock starting pressure -----------------------------------
   No counterpart in the reference: twoshock.F:141-254 always starts Newton
   from the acoustic estimate p_ac.  These entries serve callers that want
   the learned start: a 6 -> 32 -> 32 -> 1 FP32 tanh network predicts the
   correction c = log(p* / p_ac) per interface (so log(p_ac / p_left) + c
   predicts log(p* / p_left)); Newton starts from p_ac e^c where the
   acoustic wave strength passes the gate g (|p_ac - p_l| > g p_l or
   |p_ac - p_r| > g p_r), from p_ac elsewhere, runs unchanged to the same
   tolerance, and restarts from p_ac when it does not converge.  Weights are
   hipPPMLearnParameterCount() floats: W1[32][6] b1[32] W2[32][32] b2[32]
   W3[32] b3, input mean[6], 1/std[6], gate.

   hipPPMSetLearnedGuess: use the network from the next sweep on (copied;
   weights NULL switches back to the reference start).  Every weight must be
   finite, except the gate, which may be +infinity: no interface passes,
   and the plain kernels run (the reference start, at no cost). */
int hipPPMSetLearnedGuess(hipPPMData *Data, const float *weights);

/* Harvest of interface states: the sweeps record, for one owned interface in
   `keep` (by a hash of its position), the two-shock inputs and result as
   10 doubles {rho_l, rho_r, p_l, p_r, u_l, u_r, p*, starting pressure,
   Newton updates, restarted}, up to `capacity` records (0 switches it off).
   hipPPMGetHarvest copies up to max_records of them to the host, returns how
   many interfaces were selected (`seen`, may exceed the capacity) and how
   many were copied (`stored`), and restarts the count. */
int hipPPMSetHarvest(hipPPMData *Data, enzo_long_int capacity, enzo_int keep);
int hipPPMGetHarvest(hipPPMData *Data, enzo_float *records, enzo_long_int max_records,
                     enzo_long_int *seen, enzo_long_int *stored);

/* Training on the device (FP32, mean squared error of the correction
   log(p* / p_ac) clipped to [-8, 8], minibatch Adam with the learning rate
   decayed linearly from lr to lr / 10) on the records with
   positive states, p* and p_ac: `epochs` passes of shuffled minibatches of
   `batch` (>= 256) records.  The input standardisation is set from the
   records; `weights` holds the initial parameters on entry and the trained
   ones on return (the gate is left as given).  mse_per_epoch (optional,
   `epochs` values) receives the loss after each pass. */
int hipPPMLearnTrain(const enzo_float *records, enzo_long_int n, enzo_float gamma, enzo_int epochs,
                     enzo_int batch, enzo_float lr, enzo_int seed, float *weights,
                     enzo_float *mse_per_epoch);

/* Full-batch gradient (hipPPMLearnParameterCount() - 13 values, the trained
   parameters) and mean squared error over the same records as training,
   computed by the training kernels. */
int hipPPMLearnGradient(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                        const float *weights, enzo_float *grad, enzo_float *mse);

/* Network outputs (the correction c) for all n records, and their mean
   squared error over the records training would use. */
int hipPPMLearnPredict(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                       const float *weights, float *predictions, enzo_float *mse);

/* Per record, out[7] = {Newton updates from the reference start, updates
   from the learned start (restart included), restarted, p* from the
   reference start, p* from the learned start, the learned start, the
   reference start}. */
int hipPPMLearnNewtonStats(const enzo_float *records, enzo_long_int n, enzo_float gamma,
                           enzo_float pmin, const float *weights, enzo_float *out);

int hipPPMLearnParameterCount(void);
const char *hipPPMLearnLastError(void);

/* Library version string and device count check. */
const char *hipPPMVersion(void);

/* Text of the last HIP failure seen by this thread ("" if none). */
const char *hipPPMLastError(void);
int hipPPMDeviceCount(void);

/* ---- batched update of one AMR level --------------------------------------
   No counterpart entry in the reference: Enzo updates a level grid by grid
   (EvolveLevel.C's loop over grids -> grid::SolveHydroEquations ->
   grid::SolvePPM_DE, Grid_SolveHydroEquations.C:513-517).  These entries do
   that loop's work for all grids of a level at once -- each grid bitwise as
   its own hipPPMSolvePPM_DE with the level's dt would -- with the grids
   resident in one device arena, one launch per tile shape per sweep over all
   grids, and each step's launches replayed from a captured hipGraph.
   Subgrid face fluxes are not captured on this path. */
typedef struct hipPPMLevel hipPPMLevel;

/* Para: NumberOfGrids parameter blocks, one per grid (own GridDimension,
   GridStartIndex/EndIndex, dx / CellWidth), with identical hydro switches
   and field numbers.  Every axis below GridRank needs more than one active
   cell.  stream: hipStream_t or NULL (the level then makes its own, which
   stream capture requires). */
int hipPPMLevelCreate(hipPPMLevel **Level, enzo_int NumberOfGrids,
                      const hipPPMParameter *Para, void *stream);

/* Host -> device (and back) for grid `grid` of the level; synchronous. */
int hipPPMLevelSetBaryon(hipPPMLevel *Level, enzo_int grid, enzo_float *BaryonField[],
                         enzo_float *AccelerationField[]);
int hipPPMLevelGetBaryon(hipPPMLevel *Level, enzo_int grid, enzo_float *BaryonField[]);

/* One level step, queued on the level's stream: the stale pressure of every
   grid (Grid_SolvePPM_DE.C:58-61), then -- with dtFixed <= 0 -- the level's
   Courant step cfl * min over the grids of calc_dt (calc_dt.F:129-177), and
   the x/y/z sweeps of every grid in the rotation of CycleNumber. */
int hipPPMLevelStep(hipPPMLevel *Level, enzo_int CycleNumber, enzo_float dtFixed,
                    enzo_float cfl);

/* Face list of one grid's subgrids (SubgridFluxes of
   Grid_SolveHydroEquations.C:320-369 passed to the sweeps of Grid_SolvePPM_DE.C:88-143):
   from the next hipPPMLevelStep on, every sweep of that grid records the
   fluxes of those faces (Grid_xEulerSweep.C:313-376).  Waits for the level's
   stream; NumberOfSubgrids = 0 switches the grid's capture off. */
int hipPPMLevelSetSubgridFluxes(hipPPMLevel *Level, enzo_int grid,
                                enzo_fluxes *SubgridFluxes[], enzo_int NumberOfSubgrids,
                                const enzo_long_int GridGlobalStart[]);

/* The fluxes of the last hipPPMLevelStep along every axis into the
   SubgridFluxes arrays registered for the grid (the per-grid loop of
   cuPPMSaveSubgridFluxes, cuPPM.h:101-103, over x, y and z).  Waits for
   the level's stream. */
int hipPPMLevelSaveSubgridFluxes(hipPPMLevel *Level, enzo_int grid,
                                 enzo_fluxes *SubgridFluxes[], enzo_int NumberOfSubgrids);

/* hipPPMCorrectForRefinedFluxes on grid `grid` of the level, in place on
   the level's device state (grid::CorrectForRefinedFluxes,
   Grid_CorrectForRefinedFluxes.C, as UpdateFromFinerGrids.C:150-198 calls
   it): same arguments and rules.  Waits for the level's stream. */
int hipPPMLevelCorrectForRefinedFluxes(hipPPMLevel *Level, enzo_int grid,
                                       const enzo_fluxes *InitialFluxes,
                                       const enzo_fluxes *RefinedFluxes,
                                       enzo_fluxes *BoundaryFluxesThisTimeStep,
                                       enzo_int SUBlingGrid, const enzo_int *GlobalDim,
                                       const enzo_int *PeriodicFace,
                                       enzo_int CorrectParentBoundaryFlux,
                                       const enzo_int *ColourKind);

/* Waits for the level's stream; returns ENZO_PPM_SUCCESS or the error of
   any step since the last call, the dt of the last step and the fallback
   cells since the last call (the counts then restart). */
int hipPPMLevelStatus(hipPPMLevel *Level, enzo_float *dt_used, enzo_long_int *fallback_cells);

/* Replay captured hipGraphs (default) or launch each step directly. */
int hipPPMLevelUseGraph(hipPPMLevel *Level, enzo_int enable);

/* Sweep launches: enable 1 = persistent (a few workgroups per CU pull the
   level's work items from per-XCD counters until the list is done), 0 = one
   workgroup per work item (default); any other value leaves the setting.
   Returns the setting (0 or 1).  hipPPMLevelTune times both and keeps the
   faster.  Results are bitwise the same either way. */
int hipPPMLevelPersistent(hipPPMLevel *Level, enzo_int enable);

/* Tile shape of grid `grid` along sweep axis `dim`: 0 = 16 pencils x 16
   interfaces, 1 = 8 x 32 (default), 2 = 4 x 64; hipPPMLevelCommitShapes
   rebuilds the work lists. */
int hipPPMLevelSetShape(hipPPMLevel *Level, enzo_int grid, enzo_int dim, enzo_int shape);
int hipPPMLevelCommitShapes(hipPPMLevel *Level);

/* Autotuner: chooses a tile shape per pencil-length class ([0,20], [21,40],
   [41,80], longer; ghosts included) by timing `repeats` level steps of
   length dt for each choice; the level's state is restored afterwards.
   shape_by_length[4] receives the choices. */
int hipPPMLevelTune(hipPPMLevel *Level, enzo_int CycleNumber, enzo_float dt, enzo_int repeats,
                    enzo_int *shape_by_length);

/* What the last hipPPMLevelTune measured, seconds per level step:
   times[3 * class + shape] for the four length classes and the three tile
   shapes (NaN where a shape was not timed: no grid of the class, or a shape
   that cannot tile it), then times[12] for one workgroup per work item and
   times[13] for the persistent launch (NaN before any tuning). */
int hipPPMLevelTuneTimes(const hipPPMLevel *Level, enzo_float *times);

/* Comoving coordinates for the level, as hipPPMSetComoving /
   hipPPMSetExpansionFactors / hipPPMComovingExpansionTerms for one grid:
   with comoving on, every hipPPMLevelStep first keeps the state of each
   grid.  hipPPMLevelTimeStep computes the level's Courant step (pressure
   and calc_dt with a_courant) and returns it, so that the caller can take
   a(t + dt/2) before hipPPMLevelStep(..., dt, ...). */
int hipPPMLevelSetComoving(hipPPMLevel *Level, enzo_int enable);
int hipPPMLevelSetExpansionFactors(hipPPMLevel *Level, enzo_float a_courant, enzo_float a_sweep);
int hipPPMLevelTimeStep(hipPPMLevel *Level, enzo_float CourantSafetyNumber, enzo_float *dt);
int hipPPMLevelComovingExpansionTerms(hipPPMLevel *Level, enzo_float coef, enzo_float Coefficient);

/* Periodic ghost fill of every grid of the level (for grids that span a
   periodic domain alone, as a root grid does: the copies from their own
   periodic images that SetBoundaryConditions makes, see
   hipPPMPeriodicBoundary, whose argument checks apply to every grid). */
int hipPPMLevelPeriodicBoundary(hipPPMLevel *Level);

/* hipPPMExternalBoundary for every grid of the level
   (SetBoundaryConditions.C:112-113 on level 0): GridOffset holds 3 entries
   per grid; a grid's faces on the domain's edge are filled, in the
   reference's face order. */
int hipPPMLevelExternalBoundary(hipPPMLevel *Level, const enzo_int *BoundaryType,
                                const enzo_int *GridOffset, const enzo_int *BoundaryDimension,
                                enzo_int NumberOfBaryonFields, const enzo_float *InflowValue);

/* Sibling ghost copies (grid::CopyZonesFromGrid between the grids of one
   level, as SetBoundaryConditions applies it; Grid_CopyZonesFromGrid.C:
   43-230 without shearing).  hipPPMLevelSetSiblings records, from each
   grid's GridGlobalStart (3 per grid: the global index of its first cell,
   ghosts included, as SolvePPM_DE receives it) and the domain's active
   cells per axis, every overlap of a grid's whole extent with another
   grid's active region -- with the domain's periodic images when
   `periodic` (a single grid spanning the domain then gets exactly the
   periodic fill).  hipPPMLevelCopyZones applies them to the current state,
   every field the level holds (those the hydro step uses: no Vel3 in 2-D,
   no Vel2/Vel3 in 1-D, no gas energy without dual energy), in one launch. */
int hipPPMLevelSetSiblings(hipPPMLevel *Level, const enzo_long_int *GridGlobalStart,
                           const enzo_long_int *DomainActive, enzo_int periodic);
int hipPPMLevelCopyZones(hipPPMLevel *Level);

/* The learned two-shock start for every grid of the level (see
   hipPPMSetLearnedGuess; NULL switches back to the reference start). */
int hipPPMLevelSetLearnedGuess(hipPPMLevel *Level, const float *weights);

int hipPPMLevelDestroy(hipPPMLevel *Level);

/* Text of the last HIP failure of a level call on this thread. */
const char *hipPPMLevelLastError(void);

#ifdef __cplusplus
}
#endif
#endif /* ENZO_PPM_H */
