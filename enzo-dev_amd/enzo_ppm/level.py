"""Batched AMR-level update through the C ABI (hipPPMLevel*, include/enzo_ppm.h).

A Level holds every grid of one AMR level on the device and advances them
together; it computes, grid for grid, what a loop of hipPPMSolvePPM_DE calls
with the level's dt computes.  There is no CPU path: calls raise when the
library is missing or a call fails.
"""
import ctypes as C

import numpy as np

from . import abi
from .grid import make_globals


class Level:
    def __init__(self, grids, glob, lib=None, stream=None, global_starts=None):
        """global_starts: per grid, the global index of its first cell (the
        GridGlobalStart the flux correction reads); zeros by default."""
        self.lib = L = lib or abi.load_library()
        self.grids = list(grids)
        _, gd = make_globals(**glob)
        self.paras = (abi.HipPPMParameter * len(self.grids))()
        for i, g in enumerate(self.grids):
            gs = (0, 0, 0) if global_starts is None else tuple(global_starts[i])
            self.paras[i] = g.hip_parameter(L, gd, global_start=gs)
        self.handle = C.c_void_p(None)
        self._check(L.hipPPMLevelCreate(C.byref(self.handle), len(self.grids), self.paras, stream),
                    "LevelCreate")
        self.upload()

    def _check(self, st, what):
        if st != abi.SUCCESS:
            raise RuntimeError(f"{what} failed with status {st}: "
                               f"{self.lib.hipPPMLevelLastError().decode()}")

    def upload(self):
        for i, g in enumerate(self.grids):
            bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
            for k, f in enumerate(g.fields):
                bf[k] = abi.dptr(f)
            acc = (abi.PF * 3)()
            if g.accel is not None:
                for d in range(3):
                    acc[d] = abi.dptr(g.accel[d])
            self._check(self.lib.hipPPMLevelSetBaryon(self.handle, i, bf, acc), "LevelSetBaryon")

    def download(self):
        for i, g in enumerate(self.grids):
            bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
            for k, f in enumerate(g.fields):
                bf[k] = abi.dptr(f)
            self._check(self.lib.hipPPMLevelGetBaryon(self.handle, i, bf), "LevelGetBaryon")

    def step(self, cycle, dt=0.0, cfl=0.4):
        """One level step; dt <= 0 uses the level's Courant step."""
        self._check(self.lib.hipPPMLevelStep(self.handle, cycle, dt, cfl), "LevelStep")

    def set_subgrid_fluxes(self, grid, fluxes):
        """Record the face fluxes of `fluxes` (a grid.SubgridFluxes of grid
        `grid`) from the next step on; None switches the grid's capture off."""
        if fluxes is None:
            self._check(self.lib.hipPPMLevelSetSubgridFluxes(self.handle, grid, None, 0, None),
                        "LevelSetSubgridFluxes")
            return
        self._check(self.lib.hipPPMLevelSetSubgridFluxes(self.handle, grid, fluxes.ptrs, fluxes.n,
                                                         fluxes.global_start),
                    "LevelSetSubgridFluxes")

    def save_subgrid_fluxes(self, grid, fluxes):
        """The fluxes of the last step into the arrays of `fluxes`."""
        self._check(self.lib.hipPPMLevelSaveSubgridFluxes(self.handle, grid, fluxes.ptrs, fluxes.n),
                    "LevelSaveSubgridFluxes")

    def correct_for_refined_fluxes(self, grid, initial, refined, boundary=None, colour_kind=(),
                                   correct_parent_boundary_flux=0, subling=0, global_dim=None,
                                   periodic_face=None):
        """Flux correction of grid `grid` on the level's device state
        (hipPPMLevelCorrectForRefinedFluxes); arguments as
        DeviceGrid.correct_for_refined_fluxes."""
        kinds = abi.iarr(list(colour_kind), max(1, len(colour_kind)))
        self._check(self.lib.hipPPMLevelCorrectForRefinedFluxes(
            self.handle, grid, C.byref(initial), C.byref(refined),
            C.byref(boundary) if boundary is not None else None, int(subling),
            abi.iarr(global_dim, 3) if global_dim is not None else None,
            abi.iarr(periodic_face, 6) if periodic_face is not None else None,
            int(correct_parent_boundary_flux), kinds), "LevelCorrectForRefinedFluxes")

    def status(self):
        """(status, dt of the last step, fallback cells since the last call)."""
        dt = C.c_double(0.0)
        n = C.c_int64(0)
        st = self.lib.hipPPMLevelStatus(self.handle, C.byref(dt), C.byref(n))
        return st, dt.value, n.value

    def use_graph(self, on=True):
        self._check(self.lib.hipPPMLevelUseGraph(self.handle, int(on)), "LevelUseGraph")

    def persistent(self, enable=None):
        """Persistent sweep launches on (True) or off (False); None only
        queries.  Returns the setting."""
        v = -1 if enable is None else int(bool(enable))
        st = self.lib.hipPPMLevelPersistent(self.handle, v)
        if st < 0:
            raise RuntimeError(f"LevelPersistent failed with status {st}")
        return bool(st)

    def set_shape(self, grid, dim, shape):
        self._check(self.lib.hipPPMLevelSetShape(self.handle, grid, dim, shape), "LevelSetShape")

    def commit_shapes(self):
        self._check(self.lib.hipPPMLevelCommitShapes(self.handle), "LevelCommitShapes")

    def tune(self, cycle, dt, repeats=2):
        out = (abi.I * 4)()
        self._check(self.lib.hipPPMLevelTune(self.handle, cycle, dt, repeats, out), "LevelTune")
        return list(out)

    def tune_times(self):
        """The last tune's seconds per level step: ([class][shape] array,
        NaN where not timed; (ordinary, persistent) launch)."""
        import numpy as np
        t = np.full(14, np.nan)
        self._check(self.lib.hipPPMLevelTuneTimes(self.handle, abi.dptr(t)), "LevelTuneTimes")
        return t[:12].reshape(4, 3), (float(t[12]), float(t[13]))

    def set_comoving(self, enable=True):
        self._check(self.lib.hipPPMLevelSetComoving(self.handle, int(enable)), "LevelSetComoving")
        self.a_sweep = 1.0

    def set_expansion_factors(self, a_courant, a_sweep):
        self._check(self.lib.hipPPMLevelSetExpansionFactors(self.handle, a_courant, a_sweep),
                    "LevelSetExpansionFactors")
        self.a_sweep = a_sweep

    def time_step(self, cfl):
        """The level's Courant step (device), returned to the host."""
        dt = C.c_double(0.0)
        self._check(self.lib.hipPPMLevelTimeStep(self.handle, cfl, C.byref(dt)), "LevelTimeStep")
        return dt.value

    def comoving_expansion_terms(self, coef, coefficient):
        self._check(self.lib.hipPPMLevelComovingExpansionTerms(self.handle, coef, coefficient),
                    "LevelComovingExpansionTerms")

    def periodic_boundary(self):
        self._check(self.lib.hipPPMLevelPeriodicBoundary(self.handle), "LevelPeriodicBoundary")

    def external_boundary(self, types, grid_offsets, boundary_dims, inflow=None):
        """Domain-face ghost fill of every grid (hipPPMLevelExternalBoundary):
        types[2 dim + face] as in DeviceGrid.external_boundary, grid_offsets
        3 per grid (active cells from the domain's left edge),
        boundary_dims the domain's cells with ghosts, inflow[field][6]."""
        nf = self.grids[0].nfields
        vals = None
        if inflow is not None:
            self._inflow = np.ascontiguousarray(np.asarray(inflow, dtype=np.float64).reshape(nf, 6))
            vals = self._inflow.ctypes.data_as(abi.PF)
        offs = abi.iarr([v for o in grid_offsets for v in o])
        self._check(self.lib.hipPPMLevelExternalBoundary(self.handle, abi.iarr(types, 6), offs,
                                                         abi.iarr(boundary_dims, 3), nf, vals),
                    "LevelExternalBoundary")

    def set_siblings(self, global_starts, domain_active, periodic=True):
        """Sibling copy boxes from each grid's global start (index of its
        first cell, ghosts included) and the domain's active cells."""
        starts = abi.larr([v for s in global_starts for v in s])
        self._check(self.lib.hipPPMLevelSetSiblings(self.handle, starts, abi.larr(domain_active, 3),
                                                    int(periodic)), "LevelSetSiblings")

    def copy_zones(self):
        """Apply the sibling copies to the level's current state."""
        self._check(self.lib.hipPPMLevelCopyZones(self.handle), "LevelCopyZones")

    def set_learned(self, weights):
        """Learned two-shock start for every grid (None: reference start)."""
        if weights is None:
            self._check(self.lib.hipPPMLevelSetLearnedGuess(self.handle, None), "LevelSetLearned")
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        self._check(self.lib.hipPPMLevelSetLearnedGuess(self.handle, w.ctypes.data_as(C.c_void_p)),
                    "LevelSetLearned")

    def close(self):
        if self.handle.value:
            self.lib.hipPPMLevelDestroy(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
