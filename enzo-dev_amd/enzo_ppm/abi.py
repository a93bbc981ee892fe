"""ctypes mirror of include/enzo_ppm.h (the C ABI of the HIP PPM library).

The struct layouts follow the reference types they stand in for:
``enzo_fluxes`` is ``struct fluxes`` (src/enzo/Fluxes.h:29-37) with
``long_int`` = int64 and ``float`` = double (precision-64 build,
src/enzo/macros_and_parameters.h:216-217,257).
"""
import ctypes as C
import os

import numpy as np

MAX_DIMENSION = 3
MAX_NUMBER_OF_BARYON_FIELDS = 30  # PARAMETER_MAX_BARYONS, Make.config.settings:50
MAX_COLOR = 45                    # fortran.def:2

SUCCESS = 1
ERR_ARGUMENT = -1
ERR_HIP = -2
ERR_NEGATIVE = -3
ERR_NOT_SUPPORTED = -4
ERR_NO_DEVICE = -5

enzo_int = C.c_int64
enzo_float = C.c_double
PF = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int64)

Idx33 = (C.c_int64 * MAX_DIMENSION) * MAX_DIMENSION
Ptr = (PF * MAX_DIMENSION) * MAX_NUMBER_OF_BARYON_FIELDS


class EnzoFluxes(C.Structure):
    """struct fluxes (Fluxes.h:29-37)."""
    _fields_ = [
        ("LeftFluxStartGlobalIndex", Idx33),
        ("LeftFluxEndGlobalIndex", Idx33),
        ("RightFluxStartGlobalIndex", Idx33),
        ("RightFluxEndGlobalIndex", Idx33),
        ("LeftFluxes", Ptr),
        ("RightFluxes", Ptr),
    ]


class EnzoPPMGlobals(C.Structure):
    """Hydro globals read on the path (global_data.h)."""
    _fields_ = [
        ("Gamma", C.c_double),
        ("PPMDiffusionParameter", C.c_int64),
        ("PPMFlatteningParameter", C.c_int64),
        ("PPMSteepeningParameter", C.c_int64),
        ("DualEnergyFormalism", C.c_int64),
        ("DualEnergyFormalismEta1", C.c_double),
        ("DualEnergyFormalismEta2", C.c_double),
        ("PressureFree", C.c_int64),
        ("tiny_number", C.c_double),
        ("RiemannSolver", C.c_int64),
        ("RiemannSolverFallback", C.c_int64),
        ("ReconstructionMethod", C.c_int64),
        ("ConservativeReconstruction", C.c_int64),
        ("PositiveReconstruction", C.c_int64),
        ("SmallRho", C.c_double),
        ("EOSType", C.c_int64),
    ]


class EnzoPPMGrid(C.Structure):
    """Grid state that grid::SolvePPM_DE reads."""
    _fields_ = [
        ("GridRank", C.c_int64),
        ("GridDimension", C.c_int64 * 3),
        ("GridStartIndex", C.c_int64 * 3),
        ("GridEndIndex", C.c_int64 * 3),
        ("BaryonField", C.POINTER(PF)),
        ("AccelerationField", C.POINTER(PF)),
        ("DensNum", C.c_int64),
        ("GENum", C.c_int64),
        ("Vel1Num", C.c_int64),
        ("Vel2Num", C.c_int64),
        ("Vel3Num", C.c_int64),
        ("TENum", C.c_int64),
        ("dtFixed", C.c_double),
    ]


class HipPPMParameter(C.Structure):
    """hipPPMParameter (cuPPMParameter, cuPPM.h:24-50, FP64/int64)."""
    _fields_ = [
        ("DensNum", C.c_int64), ("TENum", C.c_int64), ("Vel1Num", C.c_int64),
        ("Vel2Num", C.c_int64), ("Vel3Num", C.c_int64), ("GENum", C.c_int64),
        ("MaxNumberOfBaryonFields", C.c_int64),
        ("GridGlobalStart", C.c_int64 * 3),
        ("GridStartIndex", C.c_int64 * 3),
        ("GridEndIndex", C.c_int64 * 3),
        ("GridDimension", C.c_int64 * 3),
        ("GridRank", C.c_int64),
        ("dx", C.c_double * 3),
        ("CellWidth", PF * 3),
        ("Gamma", C.c_double),
        ("PPMDiffusionParameter", C.c_int64),
        ("PPMFlatteningParameter", C.c_int64),
        ("PPMSteepeningParameter", C.c_int64),
        ("ConservativeReconstruction", C.c_int64),
        ("PositiveReconstruction", C.c_int64),
        ("GravityOn", C.c_int64),
        ("DualEnergyFormalism", C.c_int64),
        ("DualEnergyFormalismEta1", C.c_double),
        ("DualEnergyFormalismEta2", C.c_double),
        ("PressureFree", C.c_int64),
        ("MinimumPressure", C.c_double),
        ("RiemannSolver", C.c_int64),
        ("RiemannSolverFallback", C.c_int64),
        ("SmallRho", C.c_double),
        ("NumberOfColours", C.c_int64),
        ("colnum", C.c_int64 * MAX_COLOR),
    ]


def dptr(a):
    """double* of a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PF)


def iarr(vals, n=None):
    n = len(vals) if n is None else n
    arr = (C.c_int64 * n)()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libenzo_ppm_hip.so"
LIB_PATH = os.path.join(PACKAGE_DIR, LIB_NAME)

# exported entry points and their prototypes (include/enzo_ppm.h)
PROTOTYPES = {
    "hipPPMInitParameter": (C.c_int, [C.POINTER(HipPPMParameter)] + [C.c_int64] * 7
                            + [PI] * 4 + [C.c_double] * 4 + [C.c_int64] * 7
                            + [C.c_double] * 2 + [C.c_int64, C.c_double] + [C.c_int64] * 3
                            + [PI, C.c_int64, C.c_double]),
    "hipPPMSetCellWidth": (C.c_int, [C.POINTER(HipPPMParameter), C.c_int64, PF]),
    "hipPPMInitData": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(HipPPMParameter), C.c_void_p]),
    "hipPPMSetBaryon": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.POINTER(PF), PF,
                                  C.POINTER(PF)]),
    "hipPPMSetSubgridFluxes": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter),
                                         C.POINTER(C.POINTER(EnzoFluxes)), C.c_int64, PI]),
    "hipPPMSweep": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.c_double, C.c_int64]),
    "hipPPMSaveSubgridFluxes": (C.c_int, [C.POINTER(C.POINTER(EnzoFluxes)), C.c_int64, PI,
                                          C.c_void_p, C.POINTER(HipPPMParameter), C.c_int64]),
    "hipPPMGetBaryon": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.POINTER(PF)]),
    "hipPPMDestroy": (C.c_int, [C.c_void_p]),
    "hipPPMFallbackCount": (C.c_int64, [C.c_void_p]),
    "hipPPMSolvePPM_DE": (C.c_int, [C.POINTER(EnzoPPMGrid), C.POINTER(EnzoPPMGlobals), C.c_int64,
                                    C.c_int64, C.POINTER(C.POINTER(EnzoFluxes)), C.POINTER(PF),
                                    PI, C.c_int64, C.c_int64, PI, C.c_double,
                                    C.POINTER(C.c_void_p), C.c_void_p]),
    "hipPPMDeviceField": (PF, [C.c_void_p, C.c_int64]),
    "hipPPMPressureAndTimeStep": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.c_double,
                                            C.c_double, PF]),
    "hipPPMStep": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.c_int64, C.c_double]),
    "hipPPMPeriodicBoundary": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter)]),
    "hipPPMProfile": (C.c_int, [C.c_void_p, C.c_int64]),
    "hipPPMProfileRead": (C.c_int, [C.c_void_p, PF, PI]),
    "hipPPMSynchronize": (C.c_int, [C.c_void_p]),
    "hipPPMVersion": (C.c_char_p, []),
    "hipPPMLastError": (C.c_char_p, []),
    "hipPPMDeviceCount": (C.c_int, []),
    # batched AMR level
    "hipPPMLevelCreate": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, C.POINTER(HipPPMParameter),
                                    C.c_void_p]),
    "hipPPMLevelSetBaryon": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(PF), C.POINTER(PF)]),
    "hipPPMLevelGetBaryon": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(PF)]),
    "hipPPMLevelStep": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_double]),
    "hipPPMLevelSetSubgridFluxes": (C.c_int, [C.c_void_p, C.c_int64,
                                              C.POINTER(C.POINTER(EnzoFluxes)), C.c_int64, PI]),
    "hipPPMLevelSaveSubgridFluxes": (C.c_int, [C.c_void_p, C.c_int64,
                                               C.POINTER(C.POINTER(EnzoFluxes)), C.c_int64]),
    "hipPPMLevelStatus": (C.c_int, [C.c_void_p, PF, PI]),
    "hipPPMLevelUseGraph": (C.c_int, [C.c_void_p, C.c_int64]),
    "hipPPMLevelSetShape": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int64]),
    "hipPPMLevelCommitShapes": (C.c_int, [C.c_void_p]),
    "hipPPMLevelTune": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_int64, PI]),
    "hipPPMLevelSetLearnedGuess": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hipPPMLevelSetComoving": (C.c_int, [C.c_void_p, C.c_int64]),
    "hipPPMLevelSetExpansionFactors": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "hipPPMLevelTimeStep": (C.c_int, [C.c_void_p, C.c_double, PF]),
    "hipPPMLevelComovingExpansionTerms": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "hipPPMLevelPeriodicBoundary": (C.c_int, [C.c_void_p]),
    "hipPPMLevelSetSiblings": (C.c_int, [C.c_void_p, PI, PI, C.c_int64]),
    "hipPPMLevelCopyZones": (C.c_int, [C.c_void_p]),
    "hipPPMLevelDestroy": (C.c_int, [C.c_void_p]),
    "hipPPMSetLearnedGuess": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hipPPMSetComoving": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.c_int64]),
    "hipPPMProjectBoundaryFluxes": (C.c_int, [C.POINTER(EnzoFluxes), C.c_int64, C.c_int64, PI,
                                              C.POINTER(EnzoFluxes)]),
    "hipPPMExternalBoundary": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), PI, PI, PI,
                                         C.c_int64, PF]),
    "hipPPMLevelExternalBoundary": (C.c_int, [C.c_void_p, PI, PI, PI, C.c_int64, PF]),
    "hipPPMCorrectForRefinedFluxes": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter),
                                                C.POINTER(EnzoFluxes), C.POINTER(EnzoFluxes),
                                                C.POINTER(EnzoFluxes), C.c_int64, PI, PI,
                                                C.c_int64, PI]),
    "hipPPMLevelCorrectForRefinedFluxes": (C.c_int, [C.c_void_p, C.c_int64,
                                                C.POINTER(EnzoFluxes), C.POINTER(EnzoFluxes),
                                                C.POINTER(EnzoFluxes), C.c_int64, PI, PI,
                                                C.c_int64, PI]),
    "hipPPMSetExpansionFactors": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter), C.c_double,
                                            C.c_double]),
    "hipPPMComovingExpansionTerms": (C.c_int, [C.c_void_p, C.POINTER(HipPPMParameter),
                                               C.c_double, C.c_double]),
    "hipPPMSetHarvest": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64]),
    "hipPPMGetHarvest": (C.c_int, [C.c_void_p, PF, C.c_int64, PI, PI]),
    "hipPPMLearnTrain": (C.c_int, [PF, C.c_int64, C.c_double, C.c_int64, C.c_int64, C.c_double,
                                   C.c_int64, C.c_void_p, PF]),
    "hipPPMLearnGradient": (C.c_int, [PF, C.c_int64, C.c_double, C.c_void_p, PF, PF]),
    "hipPPMLearnPredict": (C.c_int, [PF, C.c_int64, C.c_double, C.c_void_p, C.c_void_p, PF]),
    "hipPPMLearnNewtonStats": (C.c_int, [PF, C.c_int64, C.c_double, C.c_double, C.c_void_p, PF]),
    "hipPPMLearnParameterCount": (C.c_int, []),
    "hipPPMLearnLastError": (C.c_char_p, []),
    "hipPPMLevelLastError": (C.c_char_p, []),
}


def header_symbols(header_path=None):
    """Function names declared in include/enzo_ppm.h."""
    import re
    if header_path is None:
        header_path = os.path.join(PACKAGE_DIR, "..", "..", "include", "enzo_ppm.h")
    txt = open(header_path).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(hipPPM\w+)\s*\(", txt, re.M)))


_LIB = None


def load_library(path=None):
    """Load the HIP library.  Raises if it is missing: there is no fallback."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)  # local binding: several builds of the ABI can share a process
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _LIB = lib
    return lib
