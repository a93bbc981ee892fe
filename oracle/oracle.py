"""ORACLE -- test infrastructure only (tests/, smoke check, bench CPU baseline).

ctypes wrapper over the two CPU builds of oracle/Makefile:
  "port": this project's C restatement (liboracle_port.so, symbols orc_*)
  "ref" : the reference Fortran kernels built with flang from /root/reference
          (oracle/_ref/libppm_ref.so, symbols ref_*), same slice driver.
Neither is ever used by the product path.
"""
import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "enzo-dev_amd"))
from enzo_ppm import abi  # noqa: E402
from enzo_ppm.grid import make_globals  # noqa: E402

PATHS = {"port": os.path.join(HERE, "liboracle_port.so"),
         "ref": os.path.join(HERE, "_ref", "libppm_ref.so")}
PREFIX = {"port": "orc_", "ref": "ref_"}
_libs = {}


def build(kind="port"):
    target = {"port": "port", "ref": "ref"}[kind]
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def available(kind):
    return os.path.exists(PATHS[kind])


def lib(kind="port"):
    if kind in _libs:
        return _libs[kind]
    if abi.BITS != 64:
        # both builds take 64-bit integers (LARGE_INTS, like the Fortran
        # kernels they wrap): a process bound to the 32-bit ABI cannot call them
        raise RuntimeError("the oracle is built with 64-bit integers; this process imported "
                           "enzo_ppm.abi with ENZO_PPM_INTEGERS=32")
    if not os.path.exists(PATHS[kind]):
        raise RuntimeError(f"oracle library {PATHS[kind]} not built")
    L = C.CDLL(PATHS[kind])
    p = PREFIX[kind]
    f = getattr(L, p + "solve_ppm_de")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.EnzoPPMGlobals), C.c_int64, C.c_int64,
                  C.POINTER(C.POINTER(abi.EnzoFluxes)), C.POINTER(abi.PF), abi.PL, C.c_int64,
                  C.c_int64, abi.PI, C.c_double, C.c_int64, C.POINTER(C.c_int64)]
    f = getattr(L, p + "sweep_direction")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.EnzoPPMGlobals), C.c_int64, C.c_int64,
                  C.POINTER(C.POINTER(abi.EnzoFluxes)), C.POINTER(abi.PF), abi.PL, C.c_int64,
                  C.c_int64, abi.PI, abi.PF, C.c_int64, C.POINTER(C.c_int64)]
    f = getattr(L, p + "compute_pressure")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.EnzoPPMGlobals), C.c_double, abi.PF]
    f = getattr(L, p + "compute_time_step")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.EnzoPPMGlobals), C.POINTER(abi.PF),
                  C.c_double, C.c_double, C.POINTER(C.c_double)]
    f = getattr(L, p + "compute_pressure_at")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.PF), C.POINTER(abi.EnzoPPMGlobals),
                  C.c_double, C.c_double, abi.PF]
    f = getattr(L, p + "comoving_expansion_terms")
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), C.POINTER(abi.PF), C.POINTER(abi.EnzoPPMGlobals),
                  C.c_double, C.c_double]
    _libs[kind] = L
    return L


def solve_ppm_de(kind, grid, glob, cycle=0, fluxes=None, nthreads=8, min_support=0.0):
    """grid::SolvePPM_DE on `grid` (modified in place).  Returns (status, fallbacks)."""
    L = lib(kind)
    gs, glob_dict = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    cw = grid.ctypes_cellwidth()
    colnum = abi.iarr(grid.colnum, max(1, len(grid.colnum)))
    nsub = fluxes.n if fluxes is not None else 0
    fptr = fluxes.ptrs if fluxes is not None else (C.POINTER(abi.EnzoFluxes) * 1)()
    gstart = fluxes.global_start if fluxes is not None else abi.larr([0, 0, 0])
    fb = C.c_int64(0)
    st = getattr(L, PREFIX[kind] + "solve_ppm_de")(
        C.byref(cg), C.byref(gs), cycle, nsub, fptr, cw, gstart, int(grid.gravity),
        grid.ncolours, colnum, min_support, nthreads, C.byref(fb))
    return st, fb.value


def sweep_direction(kind, grid, glob, dim, pressure, fluxes=None, nthreads=8):
    """One directional sweep of every slice along `dim` on `grid` (in place),
    with `pressure` (an array of the grid's shape) as the step's stale
    pressure.  Returns (status, fallbacks)."""
    import numpy as np
    L = lib(kind)
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    cw = grid.ctypes_cellwidth()
    colnum = abi.iarr(grid.colnum, max(1, len(grid.colnum)))
    nsub = fluxes.n if fluxes is not None else 0
    fptr = fluxes.ptrs if fluxes is not None else (C.POINTER(abi.EnzoFluxes) * 1)()
    gstart = fluxes.global_start if fluxes is not None else abi.larr([0, 0, 0])
    p = np.ascontiguousarray(pressure, dtype=np.float64)
    assert p.shape == grid.shape
    fb = C.c_int64(0)
    st = getattr(L, PREFIX[kind] + "sweep_direction")(
        C.byref(cg), C.byref(gs), dim, nsub, fptr, cw, gstart, int(grid.gravity), grid.ncolours,
        colnum, abi.dptr(p), nthreads, C.byref(fb))
    return st, fb.value


def compute_pressure(kind, grid, glob, min_support=0.0):
    import numpy as np
    L = lib(kind)
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    p = np.zeros(grid.shape)
    getattr(L, PREFIX[kind] + "compute_pressure")(C.byref(cg), C.byref(gs), min_support,
                                                 abi.dptr(p))
    return p


def compute_time_step(kind, grid, glob, cfl, aye=1.0, cellwidth=None):
    """Courant step (calc_dt x CFL) with expansion factor `aye`; `cellwidth`
    (3 arrays) replaces the grid's widths when given."""
    L = lib(kind)
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    if cellwidth is None:
        cw = grid.ctypes_cellwidth()
    else:
        keep_cw = [abi.dptr(c) for c in cellwidth]  # noqa: F841
        cw = (abi.PF * 3)(*keep_cw)
    dt = C.c_double(0)
    st = getattr(L, PREFIX[kind] + "compute_time_step")(C.byref(cg), C.byref(gs), cw, cfl, aye,
                                                       C.byref(dt))
    if st != abi.SUCCESS:
        raise RuntimeError(f"compute_time_step failed {st}")
    return dt.value


def _old_fields(grid, old):
    arr = (abi.PF * max(1, len(old.fields)))()
    for k, f in enumerate(old.fields):
        arr[k] = abi.dptr(f)
    return arr


def compute_pressure_at(kind, grid, old, glob, coef, min_support=0.0):
    """Pressure of coef * grid + (1 - coef) * old (Grid_ComputePressure.C's
    interpolated branch)."""
    import numpy as np
    L = lib(kind)
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    p = np.zeros(grid.shape)
    getattr(L, PREFIX[kind] + "compute_pressure_at")(C.byref(cg), _old_fields(grid, old),
                                                    C.byref(gs), coef, min_support, abi.dptr(p))
    return p


def comoving_expansion_terms(kind, grid, old, glob, coef, coefficient):
    """grid::ComovingExpansionTerms on `grid` (in place): pressure at weight
    `coef` between `old` (the state before the step; None: no old state) and
    `grid`, then expand_terms with Coefficient = dt * (da/dt) / a."""
    L = lib(kind)
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    ob = _old_fields(grid, old) if old is not None else None
    st = getattr(L, PREFIX[kind] + "comoving_expansion_terms")(C.byref(cg), ob, C.byref(gs), coef,
                                                               coefficient)
    if st != abi.SUCCESS:
        raise RuntimeError(f"comoving_expansion_terms failed {st}")


def _fx_protos(L):
    f = L.orc_project_fluxes
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoFluxes), C.c_int64, C.c_int64, abi.PI,
                  C.POINTER(abi.EnzoFluxes)]
    f = L.orc_correct_for_refined_fluxes
    f.restype = C.c_int
    f.argtypes = [C.POINTER(abi.EnzoPPMGrid), abi.PI, C.POINTER(abi.EnzoFluxes),
                  C.POINTER(abi.EnzoFluxes), C.POINTER(abi.EnzoFluxes), C.c_int64, C.c_int64,
                  C.c_double, C.c_int64, abi.PI, abi.PI, C.c_int64, abi.PI, abi.PI]
    return L


def project_fluxes(fine, rank, nfields, refinement, projected):
    """grid::GetProjectedBoundaryFluxes: `fine` and `projected` are
    enzo_fluxes structs (projected's arrays sized for the coarse faces)."""
    L = _fx_protos(lib("port"))
    st = L.orc_project_fluxes(C.byref(fine), rank, nfields, abi.iarr(refinement, 3),
                              C.byref(projected))
    if st != abi.SUCCESS:
        raise RuntimeError(f"project_fluxes failed {st}")


def correct_for_refined_fluxes(grid, global_start, initial, refined, boundary, glob,
                               colkind, correct_parent_boundary_flux=0, subling=0,
                               global_dim=None, periodic_face=None):
    """grid::CorrectForRefinedFluxes (in place on `grid` and
    `refined`/`boundary`) for a proper subgrid, or for a SUBling with the
    level's domain size `global_dim` (3) and `periodic_face` (6: left/right
    per axis).  Returns the status."""
    L = _fx_protos(lib("port"))
    _, g = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    return L.orc_correct_for_refined_fluxes(
        C.byref(cg), abi.iarr(global_start, 3), C.byref(initial), C.byref(refined),
        C.byref(boundary) if boundary is not None else None, correct_parent_boundary_flux,
        g["DualEnergyFormalism"], g["tiny_number"], grid.ncolours,
        abi.iarr(grid.colnum, max(1, grid.ncolours)), abi.iarr(colkind, max(1, len(colkind))),
        int(subling), abi.iarr(global_dim, 3) if global_dim is not None else None,
        abi.iarr(periodic_face, 6) if periodic_face is not None else None)
