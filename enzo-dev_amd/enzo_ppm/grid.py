"""Host-side grid container mirroring the state grid::SolvePPM_DE reads.

Fields are numpy float64 arrays of shape (nz, ny, nx) (x fastest, the
reference's ``index = (k*ny + j)*nx + i``, Grid_xEulerSweep.C:86), including
ghost zones.  Field numbering follows one of the reference's layouts
(Density, TotalEnergy, Velocity1-3, GasEnergy, then colours).
"""
import ctypes as C

import numpy as np

from . import abi

# Enzo defaults (SetDefaultGlobalValues.C:57,181-195,424-426,803-813)
DEFAULT_GLOBALS = dict(
    Gamma=5.0 / 3.0,
    PPMDiffusionParameter=0,
    PPMFlatteningParameter=0,
    PPMSteepeningParameter=0,
    DualEnergyFormalism=0,
    DualEnergyFormalismEta1=1e-3,
    DualEnergyFormalismEta2=0.1,
    PressureFree=0,
    tiny_number=1e-20,
    RiemannSolver=5,          # TwoShock (typedefs.h: HLL 1, HLLC 4, TwoShock 5)
    RiemannSolverFallback=1,
    ReconstructionMethod=1,   # PPM
    ConservativeReconstruction=0,
    PositiveReconstruction=0,
    SmallRho=1e-30,
    EOSType=0,
)


def make_globals(**kw):
    g = dict(DEFAULT_GLOBALS)
    for k, v in kw.items():
        if k not in g:
            raise KeyError(k)
        g[k] = v
    s = abi.EnzoPPMGlobals()
    for k, v in g.items():
        setattr(s, k, v)
    return s, g


class Grid:
    """A uniform patch with ``ghost`` ghost zones per side (3 in Enzo)."""

    def __init__(self, active, rank=3, ghost=3, ncolours=0, dual_energy=True,
                 gravity=False, dx=None, left_edge=(0.0, 0.0, 0.0)):
        active = list(active) + [1] * (3 - len(active))
        self.rank = rank
        self.ghost = ghost
        dims, start, end = [], [], []
        for d in range(3):
            if d < rank:
                dims.append(active[d] + 2 * ghost)
                start.append(ghost)
                end.append(ghost + active[d] - 1)
            else:  # padding of Grid_SolveHydroEquations.C:380-386
                dims.append(1)
                start.append(0)
                end.append(0)
        self.dims = dims
        self.start = start
        self.end = end
        self.active = active[:3]
        self.dual_energy = bool(dual_energy)
        self.ncolours = int(ncolours)
        self.DensNum, self.TENum, self.Vel1Num, self.Vel2Num, self.Vel3Num = 0, 1, 2, 3, 4
        nf = 5
        self.GENum = nf if dual_energy else 0
        nf += 1 if dual_energy else 0
        self.colnum = list(range(nf, nf + ncolours))
        nf += ncolours
        self.nfields = nf
        shape = (dims[2], dims[1], dims[0])
        self.shape = shape
        self.fields = [np.zeros(shape, dtype=np.float64) for _ in range(nf)]
        self.gravity = bool(gravity)
        self.accel = [np.zeros(shape, dtype=np.float64) for _ in range(3)] if gravity else None
        if dx is None:
            dx = [1.0 / a for a in active[:3]]
        if np.isscalar(dx):
            dx = [dx] * 3
        self.dx = list(dx)
        self.cellwidth = [np.full(dims[d], float(self.dx[d]) if d < rank else 1.0)
                          for d in range(3)]
        self.left_edge = list(left_edge)
        self.dt = 0.0

    # -- convenience views -------------------------------------------------
    @property
    def d(self):
        return self.fields[self.DensNum]

    @property
    def e(self):
        return self.fields[self.TENum]

    @property
    def u(self):
        return self.fields[self.Vel1Num]

    @property
    def v(self):
        return self.fields[self.Vel2Num]

    @property
    def w(self):
        return self.fields[self.Vel3Num]

    @property
    def ge(self):
        return self.fields[self.GENum] if self.dual_energy else None

    def active_slice(self):
        g = self.ghost
        sl = []
        for d in (2, 1, 0):
            if d < self.rank:
                sl.append(slice(g, g + self.active[d]))
            else:
                sl.append(slice(0, 1))
        return tuple(sl)

    def cell_centers(self, d):
        """Cell-centre coordinates along axis d, ghosts included."""
        n = self.dims[d]
        g = self.ghost if d < self.rank else 0
        return self.left_edge[d] + (np.arange(n) - g + 0.5) * self.dx[d]

    def copy(self):
        o = Grid.__new__(Grid)
        o.__dict__.update(self.__dict__)
        o.fields = [f.copy() for f in self.fields]
        o.accel = [a.copy() for a in self.accel] if self.accel is not None else None
        o.cellwidth = [c.copy() for c in self.cellwidth]
        return o

    # -- boundary conditions ---------------------------------------------------
    def fill_periodic(self):
        """Periodic ghost fill of all fields (ExternalBoundary periodic case)."""
        for f in self.fields:
            periodic_fill(f, self.ghost, self.rank)

    def fill_outflow(self):
        """Transmissive ghost fill (BoundaryType 1, outflow)."""
        for f in self.fields:
            outflow_fill(f, self.ghost, self.rank)

    # -- ctypes views ----------------------------------------------------------
    def ctypes_grid(self):
        """EnzoPPMGrid pointing at this grid's arrays (keep the returned tuple alive)."""
        bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
        for i, f in enumerate(self.fields):
            assert f.flags["C_CONTIGUOUS"]
            bf[i] = abi.dptr(f)
        acc = (abi.PF * 3)()
        if self.accel is not None:
            for i in range(3):
                acc[i] = abi.dptr(self.accel[i])
        g = abi.EnzoPPMGrid()
        g.GridRank = self.rank
        for d in range(3):
            g.GridDimension[d] = self.dims[d]
            g.GridStartIndex[d] = self.start[d]
            g.GridEndIndex[d] = self.end[d]
        g.BaryonField = C.cast(bf, C.POINTER(abi.PF))
        g.AccelerationField = C.cast(acc, C.POINTER(abi.PF)) if self.accel is not None else None
        g.DensNum, g.GENum = self.DensNum, self.GENum
        g.Vel1Num, g.Vel2Num, g.Vel3Num = self.Vel1Num, self.Vel2Num, self.Vel3Num
        g.TENum = self.TENum
        g.dtFixed = self.dt
        return g, (bf, acc)

    def ctypes_cellwidth(self):
        cw = (abi.PF * 3)()
        for d in range(3):
            cw[d] = abi.dptr(self.cellwidth[d])
        return cw

    def hip_parameter(self, lib, globals_dict, gravity_on=None, global_start=(0, 0, 0)):
        """hipPPMParameter for this grid via hipPPMInitParameter + SetCellWidth."""
        G = globals_dict
        para = abi.HipPPMParameter()
        colnum = abi.iarr(self.colnum, max(1, len(self.colnum)))
        gstart = abi.larr(global_start, 3)
        st = lib.hipPPMInitParameter(
            C.byref(para), self.DensNum, self.TENum, self.Vel1Num, self.Vel2Num, self.Vel3Num,
            self.GENum, abi.MAX_NUMBER_OF_BARYON_FIELDS, gstart, abi.iarr(self.start, 3),
            abi.iarr(self.end, 3), abi.iarr(self.dims, 3), self.dx[0], self.dx[1], self.dx[2],
            G["Gamma"], G["PPMDiffusionParameter"], G["PPMFlatteningParameter"],
            G["PPMSteepeningParameter"], G["ConservativeReconstruction"],
            G["PositiveReconstruction"], int(self.gravity if gravity_on is None else gravity_on),
            G["DualEnergyFormalism"], G["DualEnergyFormalismEta1"], G["DualEnergyFormalismEta2"],
            G["PressureFree"], G["tiny_number"], G["RiemannSolver"], G["RiemannSolverFallback"],
            self.ncolours, colnum, self.rank, G["SmallRho"])
        if st != abi.SUCCESS:
            raise RuntimeError(f"hipPPMInitParameter failed: {st}")
        for d in range(3):
            st = lib.hipPPMSetCellWidth(C.byref(para), d, abi.dptr(self.cellwidth[d]))
            if st != abi.SUCCESS:
                raise RuntimeError(f"hipPPMSetCellWidth failed: {st}")
        return para


def periodic_fill(f, g, rank=3):
    """Fill ghost zones of a (nz, ny, nx) array periodically along each active axis."""
    for ax_arr, ax_rank in ((2, 0), (1, 1), (0, 2)):
        if ax_rank >= rank:
            continue
        n = f.shape[ax_arr]
        a = n - 2 * g
        idx_lo = [slice(None)] * 3
        idx_hi = [slice(None)] * 3
        src_lo = [slice(None)] * 3
        src_hi = [slice(None)] * 3
        idx_lo[ax_arr] = slice(0, g)
        src_lo[ax_arr] = slice(a, a + g)
        idx_hi[ax_arr] = slice(g + a, n)
        src_hi[ax_arr] = slice(g, 2 * g)
        f[tuple(idx_lo)] = f[tuple(src_lo)]
        f[tuple(idx_hi)] = f[tuple(src_hi)]


def outflow_fill(f, g, rank=3):
    """Fill ghost zones by copying the nearest active cell (outflow)."""
    for ax_arr, ax_rank in ((2, 0), (1, 1), (0, 2)):
        if ax_rank >= rank:
            continue
        n = f.shape[ax_arr]
        for i in range(g):
            dst = [slice(None)] * 3
            src = [slice(None)] * 3
            dst[ax_arr] = slice(i, i + 1)
            src[ax_arr] = slice(g, g + 1)
            f[tuple(dst)] = f[tuple(src)]
            dst[ax_arr] = slice(n - 1 - i, n - i)
            src[ax_arr] = slice(n - 1 - g, n - g)
            f[tuple(dst)] = f[tuple(src)]


class SubgridFluxes:
    """A list of ``struct fluxes`` with numpy-backed, zeroed flux arrays
    (allocation of Grid_SolveHydroEquations.C:320-369).

    ``boxes`` are (start, end) global cell-index triples of each subgrid's
    active region (inclusive), in the parent's index space; the flux faces are
    the subgrid's outer faces.
    """

    def __init__(self, grid, boxes, global_start=(0, 0, 0)):
        self.grid = grid
        self.structs = []
        self.arrays = []  # [n][field][dim] -> (left, right)
        nf = grid.nfields
        for (s, e) in boxes:
            F = abi.EnzoFluxes()
            arrs = {}
            for dim in range(3):
                for j in range(3):
                    F.LeftFluxStartGlobalIndex[dim][j] = s[j]
                    F.LeftFluxEndGlobalIndex[dim][j] = e[j]
                    F.RightFluxStartGlobalIndex[dim][j] = s[j]
                    F.RightFluxEndGlobalIndex[dim][j] = e[j]
                F.LeftFluxEndGlobalIndex[dim][dim] = s[dim]
                F.RightFluxStartGlobalIndex[dim][dim] = e[dim]
                size = 1
                for j in range(3):
                    size *= (F.LeftFluxEndGlobalIndex[dim][j] - F.LeftFluxStartGlobalIndex[dim][j] + 1)
                for fld in range(nf):
                    L = np.zeros(size)
                    Rr = np.zeros(size)
                    arrs[(fld, dim)] = (L, Rr)
                    F.LeftFluxes[fld][dim] = abi.dptr(L)
                    F.RightFluxes[fld][dim] = abi.dptr(Rr)
            self.structs.append(F)
            self.arrays.append(arrs)
        n = len(self.structs)
        self.ptrs = (C.POINTER(abi.EnzoFluxes) * max(1, n))()
        for i, F in enumerate(self.structs):
            self.ptrs[i] = C.pointer(F)
        self.n = n
        self.global_start = abi.larr(global_start, 3)

    def zero(self):
        for arrs in self.arrays:
            for (L, Rr) in arrs.values():
                L[:] = 0
                Rr[:] = 0

    def flat(self):
        """All flux arrays concatenated in a fixed order (for comparisons)."""
        out = []
        for arrs in self.arrays:
            for key in sorted(arrs):
                out.append(arrs[key][0])
                out.append(arrs[key][1])
        return np.concatenate(out) if out else np.zeros(0)
