"""ctypes mirror of include/enzo_ppm.h (the C ABI of the HIP PPM library).

The struct layouts follow the reference types they stand in for:
``enzo_fluxes`` is ``struct fluxes`` (src/enzo/Fluxes.h:29-37) with
``long_int`` = int64 and ``float`` = double (precision-64 build,
src/enzo/macros_and_parameters.h:151,257).  ``int`` is int64 (integers-64,
:216-217) unless the environment sets ENZO_PPM_INTEGERS=32 before this
module is imported: then it is int32 (integers-32, :201-213) and the
library is libenzo_ppm_hip_i32.so, the header's ENZO_PPM_INTEGERS_32 build.
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

BITS = int(os.environ.get("ENZO_PPM_INTEGERS", "64"))
if BITS not in (32, 64):
    raise ValueError(f"ENZO_PPM_INTEGERS must be 32 or 64, not {BITS}")
enzo_int = C.c_int32 if BITS == 32 else C.c_int64   # `int`
enzo_long_int = C.c_int64                           # long_int
enzo_float = C.c_double
I = enzo_int
LI = enzo_long_int
PF = C.POINTER(C.c_double)
PI = C.POINTER(enzo_int)
PL = C.POINTER(enzo_long_int)

Idx33 = (LI * MAX_DIMENSION) * MAX_DIMENSION
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
        ("PPMDiffusionParameter", I),
        ("PPMFlatteningParameter", I),
        ("PPMSteepeningParameter", I),
        ("DualEnergyFormalism", I),
        ("DualEnergyFormalismEta1", C.c_double),
        ("DualEnergyFormalismEta2", C.c_double),
        ("PressureFree", I),
        ("tiny_number", C.c_double),
        ("RiemannSolver", I),
        ("RiemannSolverFallback", I),
        ("ReconstructionMethod", I),
        ("ConservativeReconstruction", I),
        ("PositiveReconstruction", I),
        ("SmallRho", C.c_double),
        ("EOSType", I),
    ]


class EnzoPPMGrid(C.Structure):
    """Grid state that grid::SolvePPM_DE reads."""
    _fields_ = [
        ("GridRank", I),
        ("GridDimension", I * 3),
        ("GridStartIndex", I * 3),
        ("GridEndIndex", I * 3),
        ("BaryonField", C.POINTER(PF)),
        ("AccelerationField", C.POINTER(PF)),
        ("DensNum", I),
        ("GENum", I),
        ("Vel1Num", I),
        ("Vel2Num", I),
        ("Vel3Num", I),
        ("TENum", I),
        ("dtFixed", C.c_double),
    ]


class HipPPMParameter(C.Structure):
    """hipPPMParameter (cuPPMParameter, cuPPM.h:24-50, FP64, `int` = enzo_int)."""
    _fields_ = [
        ("DensNum", I), ("TENum", I), ("Vel1Num", I),
        ("Vel2Num", I), ("Vel3Num", I), ("GENum", I),
        ("MaxNumberOfBaryonFields", I),
        ("GridGlobalStart", LI * 3),
        ("GridStartIndex", I * 3),
        ("GridEndIndex", I * 3),
        ("GridDimension", I * 3),
        ("GridRank", I),
        ("dx", C.c_double * 3),
        ("CellWidth", PF * 3),
        ("Gamma", C.c_double),
        ("PPMDiffusionParameter", I),
        ("PPMFlatteningParameter", I),
        ("PPMSteepeningParameter", I),
        ("ConservativeReconstruction", I),
        ("PositiveReconstruction", I),
        ("GravityOn", I),
        ("DualEnergyFormalism", I),
        ("DualEnergyFormalismEta1", C.c_double),
        ("DualEnergyFormalismEta2", C.c_double),
        ("PressureFree", I),
        ("MinimumPressure", C.c_double),
        ("RiemannSolver", I),
        ("RiemannSolverFallback", I),
        ("SmallRho", C.c_double),
        ("NumberOfColours", I),
        ("colnum", I * MAX_COLOR),
    ]


def dptr(a):
    """double* of a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PF)


def iarr(vals, n=None):
    """enzo_int array (`int` arguments and members)."""
    n = len(vals) if n is None else n
    arr = (I * n)()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


def larr(vals, n=None):
    """long_int array (GridGlobalStart and the like)."""
    n = len(vals) if n is None else n
    arr = (LI * n)()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libenzo_ppm_hip_i32.so" if BITS == 32 else "libenzo_ppm_hip.so"
LIB_PATH = os.path.join(PACKAGE_DIR, LIB_NAME)

# exported entry points and their prototypes (include/enzo_ppm.h): I is
# enzo_int, LI long_int (tests/test_capi.py checks this table against the
# header's declarations)
_P = C.POINTER(HipPPMParameter)
_FL = C.POINTER(C.POINTER(EnzoFluxes))
_F1 = C.POINTER(EnzoFluxes)
_H = C.c_void_p
_D = C.c_double
PROTOTYPES = {
    "hipPPMInitParameter": (C.c_int, [_P] + [I] * 7 + [PL] + [PI] * 3 + [_D] * 4 + [I] * 7
                            + [_D] * 2 + [I, _D] + [I] * 3 + [PI, I, _D]),
    "hipPPMSetCellWidth": (C.c_int, [_P, I, PF]),
    "hipPPMInitData": (C.c_int, [C.POINTER(_H), _P, _H]),
    "hipPPMSetBaryon": (C.c_int, [_H, _P, C.POINTER(PF), PF, C.POINTER(PF)]),
    "hipPPMSetSubgridFluxes": (C.c_int, [_H, _P, _FL, I, PL]),
    "hipPPMSweep": (C.c_int, [_H, _P, _D, I]),
    "hipPPMSaveSubgridFluxes": (C.c_int, [_FL, I, PL, _H, _P, I]),
    "hipPPMGetBaryon": (C.c_int, [_H, _P, C.POINTER(PF)]),
    "hipPPMDestroy": (C.c_int, [_H]),
    "hipPPMFallbackCount": (LI, [_H]),
    "hipPPMSolvePPM_DE": (C.c_int, [C.POINTER(EnzoPPMGrid), C.POINTER(EnzoPPMGlobals), I, I, _FL,
                                    C.POINTER(PF), PL, I, I, PI, _D, C.POINTER(_H), _H]),
    "hipPPMSolvePPM_DE_NoAlloc": (C.c_int, [C.POINTER(EnzoPPMGrid), C.POINTER(EnzoPPMGlobals), I,
                                            C.POINTER(PF), I, I, PI, _D, _H, PL]),
    "hipPPMDeviceField": (PF, [_H, I]),
    "hipPPMPressureAndTimeStep": (C.c_int, [_H, _P, _D, _D, PF]),
    "hipPPMStep": (C.c_int, [_H, _P, I, _D]),
    "hipPPMPeriodicBoundary": (C.c_int, [_H, _P]),
    "hipPPMExternalBoundary": (C.c_int, [_H, _P, PI, PI, PI, I, PF]),
    "hipPPMProjectBoundaryFluxes": (C.c_int, [_F1, I, I, PI, _F1]),
    "hipPPMCorrectForRefinedFluxes": (C.c_int, [_H, _P, _F1, _F1, _F1, I, PI, PI, I, PI]),
    "hipPPMSetComoving": (C.c_int, [_H, _P, I]),
    "hipPPMSetExpansionFactors": (C.c_int, [_H, _P, _D, _D]),
    "hipPPMComovingExpansionTerms": (C.c_int, [_H, _P, _D, _D]),
    "hipPPMProfile": (C.c_int, [_H, I]),
    "hipPPMProfileRead": (C.c_int, [_H, PF, PL]),
    "hipPPMSynchronize": (C.c_int, [_H]),
    "hipPPMSetLearnedGuess": (C.c_int, [_H, C.c_void_p]),
    "hipPPMSetHarvest": (C.c_int, [_H, LI, I]),
    "hipPPMGetHarvest": (C.c_int, [_H, PF, LI, PL, PL]),
    "hipPPMLearnTrain": (C.c_int, [PF, LI, _D, I, I, _D, I, C.c_void_p, PF]),
    "hipPPMLearnGradient": (C.c_int, [PF, LI, _D, C.c_void_p, PF, PF]),
    "hipPPMLearnPredict": (C.c_int, [PF, LI, _D, C.c_void_p, C.c_void_p, PF]),
    "hipPPMLearnNewtonStats": (C.c_int, [PF, LI, _D, _D, C.c_void_p, PF]),
    "hipPPMLearnParameterCount": (C.c_int, []),
    "hipPPMLearnLastError": (C.c_char_p, []),
    "hipPPMVersion": (C.c_char_p, []),
    "hipPPMLastError": (C.c_char_p, []),
    "hipPPMDeviceCount": (C.c_int, []),
    # batched AMR level
    "hipPPMLevelCreate": (C.c_int, [C.POINTER(_H), I, _P, _H]),
    "hipPPMLevelSetBaryon": (C.c_int, [_H, I, C.POINTER(PF), C.POINTER(PF)]),
    "hipPPMLevelGetBaryon": (C.c_int, [_H, I, C.POINTER(PF)]),
    "hipPPMLevelStep": (C.c_int, [_H, I, _D, _D]),
    "hipPPMLevelSetSubgridFluxes": (C.c_int, [_H, I, _FL, I, PL]),
    "hipPPMLevelSaveSubgridFluxes": (C.c_int, [_H, I, _FL, I]),
    "hipPPMLevelCorrectForRefinedFluxes": (C.c_int, [_H, I, _F1, _F1, _F1, I, PI, PI, I, PI]),
    "hipPPMLevelStatus": (C.c_int, [_H, PF, PL]),
    "hipPPMLevelUseGraph": (C.c_int, [_H, I]),
    "hipPPMLevelPersistent": (C.c_int, [_H, I]),
    "hipPPMLevelSetShape": (C.c_int, [_H, I, I, I]),
    "hipPPMLevelCommitShapes": (C.c_int, [_H]),
    "hipPPMLevelTune": (C.c_int, [_H, I, _D, I, PI]),
    "hipPPMLevelTuneTimes": (C.c_int, [_H, PF]),
    "hipPPMLevelSetComoving": (C.c_int, [_H, I]),
    "hipPPMLevelSetExpansionFactors": (C.c_int, [_H, _D, _D]),
    "hipPPMLevelTimeStep": (C.c_int, [_H, _D, PF]),
    "hipPPMLevelComovingExpansionTerms": (C.c_int, [_H, _D, _D]),
    "hipPPMLevelPeriodicBoundary": (C.c_int, [_H]),
    "hipPPMLevelExternalBoundary": (C.c_int, [_H, PI, PI, PI, I, PF]),
    "hipPPMLevelSetSiblings": (C.c_int, [_H, PL, PL, I]),
    "hipPPMLevelCopyZones": (C.c_int, [_H]),
    "hipPPMLevelSetLearnedGuess": (C.c_int, [_H, C.c_void_p]),
    "hipPPMLevelDestroy": (C.c_int, [_H]),
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
