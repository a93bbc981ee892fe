"""Python driver over the HIP library (include/enzo_ppm.h).

Every call goes through the C ABI; if the library is missing or no GPU is
visible the calls raise -- there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import abi
from .grid import make_globals


def _check(st, what):
    if st != abi.SUCCESS:
        err = abi.load_library().hipPPMLastError().decode()
        raise RuntimeError(f"{what} failed with status {st}: {err}")


def solve_ppm_de(grid, glob, cycle=0, fluxes=None, handle=None, min_support=0.0, lib=None):
    """grid::SolvePPM_DE on the GPU through hipPPMSolvePPM_DE; updates
    `grid` in place and fills `fluxes` (a grid.SubgridFluxes) if given.
    Returns the status code (SUCCESS or a negative error)."""
    L = lib or abi.load_library()
    gs, _ = make_globals(**glob)
    cg, keep = grid.ctypes_grid()
    cw = grid.ctypes_cellwidth()
    colnum = abi.iarr(grid.colnum, max(1, len(grid.colnum)))
    nsub = fluxes.n if fluxes is not None else 0
    fptr = fluxes.ptrs if fluxes is not None else (C.POINTER(abi.EnzoFluxes) * 1)()
    gstart = fluxes.global_start if fluxes is not None else abi.larr([0, 0, 0])
    st = L.hipPPMSolvePPM_DE(C.byref(cg), C.byref(gs), cycle, nsub, fptr, cw, gstart,
                             int(grid.gravity), grid.ncolours, colnum, min_support,
                             C.byref(handle) if handle is not None else None, None)
    return st


class DeviceGrid:
    """Device-resident state of one grid for multi-step runs: upload once,
    then pressure + Courant pass and rotated sweeps stay on the GPU."""

    def __init__(self, grid, glob, lib=None, stream=None, global_start=(0, 0, 0)):
        self.lib = L = lib or abi.load_library()
        self.grid = grid
        self.glob = dict(glob)
        _, gd = make_globals(**glob)
        self.gdict = gd
        self.para = grid.hip_parameter(L, gd, global_start=global_start)
        self.handle = C.c_void_p(None)
        _check(L.hipPPMInitData(C.byref(self.handle), C.byref(self.para), stream), "InitData")
        self._keep = None
        self.a_sweep = 1.0
        self.upload()

    def upload(self):
        g = self.grid
        bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
        for i, f in enumerate(g.fields):
            bf[i] = abi.dptr(f)
        acc = (abi.PF * 3)()
        if g.accel is not None:
            for i in range(3):
                acc[i] = abi.dptr(g.accel[i])
        self._keep = (bf, acc)
        _check(self.lib.hipPPMSetBaryon(self.handle, C.byref(self.para), bf, None, acc),
               "SetBaryon")

    def download(self):
        g = self.grid
        bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
        for i, f in enumerate(g.fields):
            bf[i] = abi.dptr(f)
        return self.lib.hipPPMGetBaryon(self.handle, C.byref(self.para), bf)

    def pressure_and_dt(self, cfl, min_support=0.0, want_host=True):
        dt = C.c_double(0.0)
        _check(self.lib.hipPPMPressureAndTimeStep(self.handle, C.byref(self.para), cfl,
                                                   min_support,
                                                   C.byref(dt) if want_host else None),
               "PressureAndTimeStep")
        return dt.value

    def step(self, cycle, dt=0.0):
        """Rotated sweeps; dt <= 0 uses the device dt of pressure_and_dt."""
        _check(self.lib.hipPPMStep(self.handle, C.byref(self.para), cycle, dt), "Step")

    def periodic_boundary(self):
        _check(self.lib.hipPPMPeriodicBoundary(self.handle, C.byref(self.para)), "Boundary")

    def external_boundary(self, types, grid_offset=(0, 0, 0), boundary_dims=None, inflow=None):
        """Domain-face ghost fill (hipPPMExternalBoundary): types[2 dim +
        face] in the reference's numbering (reflecting 0, outflow 1, inflow 2,
        periodic 3, shearing 4); boundary_dims defaults to this grid's
        dimensions (the grid spans the domain); inflow[field][2 dim + face]."""
        bd = list(self.grid.dims) if boundary_dims is None else list(boundary_dims)
        nf = self.grid.nfields
        vals = None
        if inflow is not None:
            flat = np.ascontiguousarray(np.asarray(inflow, dtype=np.float64).reshape(nf, 6))
            self._inflow = flat
            vals = flat.ctypes.data_as(abi.PF)
        _check(self.lib.hipPPMExternalBoundary(self.handle, C.byref(self.para), abi.iarr(types, 6),
                                               abi.iarr(grid_offset, 3), abi.iarr(bd, 3), nf, vals),
               "ExternalBoundary")

    def synchronize(self):
        _check(self.lib.hipPPMSynchronize(self.handle), "Synchronize")

    def profile(self, enable=True):
        _check(self.lib.hipPPMProfile(self.handle, int(enable)), "Profile")

    def profile_read(self):
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        _check(self.lib.hipPPMProfileRead(self.handle, C.byref(ms), C.byref(n)), "ProfileRead")
        return ms.value, n.value

    def fallback_count(self):
        return self.lib.hipPPMFallbackCount(self.handle)

    def correct_for_refined_fluxes(self, initial, refined, boundary=None, colour_kind=(),
                                   correct_parent_boundary_flux=0, subling=0, global_dim=None,
                                   periodic_face=None):
        """Flux correction of the device state (hipPPMCorrectForRefinedFluxes)
        for one proper subgrid, or, with `subling`, for a finer grid that is
        not this grid's child but touches one of its faces (`global_dim`: the
        level's domain size in cells per axis; `periodic_face`: 6 flags, left
        and right face per axis).  Structs are abi.EnzoFluxes."""
        kinds = abi.iarr(list(colour_kind), max(1, len(colour_kind)))
        _check(self.lib.hipPPMCorrectForRefinedFluxes(
            self.handle, C.byref(self.para), C.byref(initial), C.byref(refined),
            C.byref(boundary) if boundary is not None else None, int(subling),
            abi.iarr(global_dim, 3) if global_dim is not None else None,
            abi.iarr(periodic_face, 6) if periodic_face is not None else None,
            int(correct_parent_boundary_flux), kinds), "CorrectForRefinedFluxes")

    def set_comoving(self, enable=True):
        """Comoving handling on the device state (hipPPMSetComoving)."""
        _check(self.lib.hipPPMSetComoving(self.handle, C.byref(self.para), int(enable)),
               "SetComoving")
        self.a_sweep = 1.0

    def set_expansion_factors(self, a_courant, a_sweep):
        """a(Time) for the Courant pass, a(Time + dt/2) for the sweeps."""
        _check(self.lib.hipPPMSetExpansionFactors(self.handle, C.byref(self.para), a_courant,
                                                  a_sweep), "SetExpansionFactors")
        self.a_sweep = a_sweep

    def comoving_expansion_terms(self, coef, coefficient):
        """Expansion terms after a step: pressure weight `coef` between the
        step's start and now, Coefficient = dt (da/dt) / a."""
        _check(self.lib.hipPPMComovingExpansionTerms(self.handle, C.byref(self.para), coef,
                                                     coefficient), "ComovingExpansionTerms")

    def set_learned(self, weights):
        """Learned two-shock start from the next sweep on (float32 weights of
        learn.N_FLOATS values); None restores the reference start."""
        if weights is None:
            _check(self.lib.hipPPMSetLearnedGuess(self.handle, None), "SetLearnedGuess")
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _check(self.lib.hipPPMSetLearnedGuess(self.handle, w.ctypes.data_as(C.c_void_p)),
               "SetLearnedGuess")

    def set_harvest(self, capacity, keep=1):
        """Record the two-shock inputs and result of one owned interface in
        `keep`, up to `capacity` records (0 switches the harvest off)."""
        _check(self.lib.hipPPMSetHarvest(self.handle, int(capacity), int(keep)), "SetHarvest")

    def get_harvest(self, max_records):
        """(records[n, 10] float64, interfaces selected since the last call)."""
        out = np.zeros((max(1, int(max_records)), 10))
        seen = C.c_int64(0)
        stored = C.c_int64(0)
        _check(self.lib.hipPPMGetHarvest(self.handle, abi.dptr(out), int(max_records),
                                         C.byref(seen), C.byref(stored)), "GetHarvest")
        return out[:stored.value].copy(), seen.value

    def close(self):
        if self.handle.value:
            self.lib.hipPPMDestroy(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
