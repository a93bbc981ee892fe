"""Checks of the oracle's slice driver (oracle/ppm_driver_oracle.c) that do not
go through the driver's own reading of the reference C++.

The driver is built into both oracle libraries (the C restatement and the
reference Fortran), so comparisons between those two cannot see a slip in it.
These tests pin it from outside:
  * a step is the three single sweeps in the order Grid_SolvePPM_DE.C:86-144
    states (x y z at cycle 0, y z x at 1, z x y at 2), all with the pressure
    of the step's start (:58-61); the test spells the order out itself;
  * relabelling the axes cyclically (x <- y <- z) with the velocities and
    cell widths relabelled the same way, and the cycle shifted to match,
    gives the relabelled result bit for bit (velocity permutation of
    Grid_[xyz]EulerSweep.C:81-133, the row / slice axes of flattening and
    diffusion, the face offsets of the flux capture :313-376);
  * transverse ghost pencils are swept (Grid_xEulerSweep.C:172-178): a ghost
    row of one grid ends like the same row made active in a wider grid;
  * the captured density and colour fluxes of every face position telescope
    to the mass change of the cells between the two faces of the pencil, so
    offsets, sides and the colour fluxes' division by dx (euler.F:173-177)
    are fixed by conservation, and a sweep from rest captures momentum flux
    only in the slot of the sweep's own velocity;
  * ComputePressure (Grid_ComputePressure.C:96-166) on hand-made cells whose
    pressures are exact in binary: both dual-energy branches, the velocities
    a rank counts, the tiny_number floor and the minimum-support floor.
"""
import numpy as np
import pytest

import oracle
from enzo_ppm import problems
from enzo_ppm.grid import Grid, SubgridFluxes
from helpers import bit_equal

KINDS = ["port"] + (["ref"] if oracle.available("ref") else [])
# sweep order of each cycle of a rank-3 grid, Grid_SolvePPM_DE.C:86-144
# (ixyz = CycleNumber % GridRank, then n = ixyz .. ixyz + GridRank - 1)
ORDER = {0: (0, 1, 2), 1: (1, 2, 0), 2: (2, 0, 1)}


def _state(flags=(1, 1, 0), gravity=True, ncolours=2, seed=3, active=(9, 7, 8)):
    g, glob, run = problems.random_smooth(list(active), seed=seed, ncolours=ncolours,
                                          gravity=gravity, flags=flags)
    g.fill_periodic()
    g.dt = oracle.compute_time_step("port", g, glob, run["cfl"])
    return g, glob


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("flags", [(1, 1, 0), (1, 1, 1)], ids=["steep-flat", "steep-flat-diffusion"])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_step_is_single_sweeps_in_reference_order(kind, flags, cycle):
    g, glob = _state(flags)
    a, b = g.copy(), g.copy()
    assert oracle.solve_ppm_de(kind, a, glob, cycle=cycle) == (1, 0)
    p0 = oracle.compute_pressure(kind, b, glob)
    for dim in ORDER[cycle]:
        assert oracle.sweep_direction(kind, b, glob, dim, p0) == (1, 0)
    assert all(bit_equal(x, y) for x, y in zip(a.fields, b.fields))
    # the comparison sees the order and the stale pressure: the reverse
    # order, or a pressure refreshed before each sweep, ends elsewhere
    c = g.copy()
    for dim in reversed(ORDER[cycle]):
        oracle.sweep_direction(kind, c, glob, dim, p0)
    assert not bit_equal(a.d, c.d)
    c = g.copy()
    for dim in ORDER[cycle]:
        oracle.sweep_direction(kind, c, glob, dim, oracle.compute_pressure(kind, c, glob))
    assert not bit_equal(a.d, c.d)


def relabel(g, s):
    """The grid with new axis m = old axis (m + s) % 3: fields transposed,
    velocity m' = old velocity (m' + s) % 3, widths and extents with them."""
    assert g.rank == 3
    o = g.copy()
    old = lambda m: (m + s) % 3  # noqa: E731
    axes = [2 - old(2 - q) for q in range(3)]  # array axis q holds spatial axis 2 - q
    for name in ("dims", "start", "end", "active", "dx"):
        v = getattr(g, name)
        setattr(o, name, [v[old(m)] for m in range(3)])
    o.cellwidth = [g.cellwidth[old(m)].copy() for m in range(3)]
    o.fields = [np.ascontiguousarray(np.transpose(f, axes)) for f in g.fields]
    vel = [g.Vel1Num, g.Vel2Num, g.Vel3Num]
    for m in range(3):
        o.fields[vel[m]] = np.ascontiguousarray(np.transpose(g.fields[vel[old(m)]], axes))
    if g.accel is not None:
        o.accel = [np.ascontiguousarray(np.transpose(g.accel[old(m)], axes)) for m in range(3)]
    o.shape = o.fields[0].shape
    return o


def relabel_box(box, s):
    lo, hi = box
    return (tuple(lo[(m + s) % 3] for m in range(3)), tuple(hi[(m + s) % 3] for m in range(3)))


def face_array(fl, n, field, dim, side):
    """Captured face fluxes of subgrid n as [higher transverse][lower transverse]."""
    arr = fl.arrays[n][(field, dim)][0 if side == 0 else 1]
    F = fl.structs[n]
    t = [a for a in range(3) if a != dim]
    ext = [F.RightFluxEndGlobalIndex[dim][a] - F.RightFluxStartGlobalIndex[dim][a] + 1 for a in t]
    return arr.reshape(ext[1], ext[0]), t


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("shift", [1, 2])
@pytest.mark.parametrize("flags", [(1, 1, 0), (1, 1, 1)], ids=["steep-flat", "steep-flat-diffusion"])
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_cyclic_axis_relabelling(kind, shift, flags, cycle):
    g, glob = _state(flags)
    h = relabel(g, shift)
    boxes = [((4, 5, 3), (8, 7, 6)), ((3, 3, 5), (5, 9, 9))]
    fg = SubgridFluxes(g, boxes)
    fh = SubgridFluxes(h, [relabel_box(b, shift) for b in boxes])
    assert oracle.solve_ppm_de(kind, g, glob, cycle=cycle, fluxes=fg) == (1, 0)
    # the relabelled grid sweeps the same physical axes in the same order
    assert oracle.solve_ppm_de(kind, h, glob, cycle=cycle + 3 - shift, fluxes=fh) == (1, 0)
    back = relabel(h, 3 - shift)
    assert all(bit_equal(x, y) for x, y in zip(g.fields, back.fields))
    vel = [g.Vel1Num, g.Vel2Num, g.Vel3Num]
    for n in range(len(boxes)):
        for dim in range(3):
            dim_h = (dim - shift) % 3
            for field in range(g.nfields):
                fld_h = field
                if field in vel:
                    fld_h = vel[(vel.index(field) - shift) % 3]
                for side in (0, 1):
                    a, ta = face_array(fg, n, field, dim, side)
                    b, tb = face_array(fh, n, fld_h, dim_h, side)
                    # tb holds the relabelled transverse axes; map them back
                    if [(x + shift) % 3 for x in tb] != ta:
                        b = b.T
                    assert bit_equal(a, b), (n, dim, field, side)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_transverse_ghost_pencils_are_swept(kind, dim):
    """Flattening (line 1 of each slice) and diffusion (neighbour pencils)
    off, so every pencil is swept on its own data: grid h has the row axis
    of g widened by three rows on the low side, which makes g's low ghost
    rows active in h; each of g's rows must end as the same row of h."""
    g, glob = _state(flags=(1, 0, 0))
    rng = np.random.default_rng(5)
    for f in g.fields:  # ghosts that no fill would produce
        f *= 1.0 + 0.01 * rng.standard_normal(f.shape)
    row = (dim + 1) % 3
    q = 2 - row  # array axis of the row axis
    h = g.copy()
    pad = [(0, 0)] * 3
    pad[q] = (3, 0)
    h.fields = [np.pad(f, pad, mode="edge").copy() for f in g.fields]
    if g.accel is not None:
        h.accel = [np.pad(a, pad, mode="edge").copy() for a in g.accel]
    h.dims = list(g.dims)
    h.dims[row] += 3
    h.end = list(g.end)
    h.end[row] += 3
    h.start = list(g.start)
    h.cellwidth = [c.copy() for c in g.cellwidth]
    h.cellwidth[row] = np.concatenate([np.full(3, g.cellwidth[row][0]), g.cellwidth[row]])
    h.shape = h.fields[0].shape
    assert oracle.sweep_direction(kind, g, glob, dim, oracle.compute_pressure(kind, g, glob))[0] == 1
    assert oracle.sweep_direction(kind, h, glob, dim, oracle.compute_pressure(kind, h, glob))[0] == 1
    sl = [slice(None)] * 3
    sl[q] = slice(3, None)
    for x, y in zip(g.fields, h.fields):
        assert bit_equal(x, y[tuple(sl)])


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_captured_fluxes_telescope_to_mass_change(kind, dim):
    g, glob = _state(flags=(1, 1, 0), ncolours=2)
    boxes = [((4, 5, 3), (8, 7, 6)), ((3, 3, 5), (5, 9, 9))]
    fl = SubgridFluxes(g, boxes)
    before = [f.copy() for f in g.fields]
    assert oracle.sweep_direction(kind, g, glob, dim, oracle.compute_pressure(kind, g, glob),
                                  fluxes=fl) == (1, 0)
    largest = 0.0
    for n, (lo, hi) in enumerate(boxes):
        t = [a for a in range(3) if a != dim]
        for field in [g.DensNum] + g.colnum:
            L, _ = face_array(fl, n, field, dim, 0)
            R, _ = face_array(fl, n, field, dim, 1)
            change = g.fields[field] - before[field]
            for jh in range(hi[t[1]] - lo[t[1]] + 1):
                for jl in range(hi[t[0]] - lo[t[0]] + 1):
                    idx = [0, 0, 0]
                    idx[t[0]] = lo[t[0]] + jl
                    idx[t[1]] = lo[t[1]] + jh
                    cells = []
                    for c in range(lo[dim], hi[dim] + 1):
                        idx[dim] = c
                        cells.append(change[idx[2], idx[1], idx[0]])
                    total = float(np.sum(cells))
                    flux = L[jh, jl] - R[jh, jl]
                    scale = np.max(np.abs(before[field])) * len(cells)
                    assert abs(total - flux) <= 4e-15 * scale, (n, field, jh, jl, total, flux)
                    largest = max(largest, abs(flux) / scale)
    assert largest > 1e-4  # the faces do carry mass


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_sweep_from_rest_moves_only_its_own_momentum(kind, dim):
    g, glob = _state(flags=(1, 1, 0), gravity=False, ncolours=1)
    for v in (g.u, g.v, g.w):
        v[...] = 0.0
    e = g.fields[g.GENum].copy()
    g.fields[g.TENum][...] = e  # at rest total = thermal energy
    boxes = [((4, 5, 3), (8, 7, 6))]
    fl = SubgridFluxes(g, boxes)
    assert oracle.sweep_direction(kind, g, glob, dim, oracle.compute_pressure(kind, g, glob),
                                  fluxes=fl) == (1, 0)
    vel = [g.Vel1Num, g.Vel2Num, g.Vel3Num]
    for m in range(3):
        moved = np.any(g.fields[vel[m]] != 0.0)
        assert moved == (m == dim), m
        for side in (0, 1):
            arr, _ = face_array(fl, 0, vel[m], dim, side)
            assert np.any(arr != 0.0) == (m == dim), (m, side)


def _cells(values, rank=3, dual=True):
    """A 1-active-cell-per-axis grid whose cells carry `values` as
    (density, total energy, vx, vy, vz, gas energy) along x."""
    n = len(values)
    g = Grid([n, 1, 1][:rank] if rank > 1 else [n], rank=rank, dual_energy=dual, ncolours=0)
    for i, (d, e, vx, vy, vz, ge) in enumerate(values):
        idx = (slice(None), slice(None), 3 + i)
        g.d[idx], g.e[idx], g.u[idx] = d, e, vx
        if rank > 1:
            g.v[idx] = vy
        if rank > 2:
            g.w[idx] = vz
        if dual:
            g.fields[g.GENum][idx] = ge
    return g


@pytest.mark.parametrize("kind", KINDS)
def test_compute_pressure_closed_form(kind):
    """Pressures on hand-made cells, with gamma = 3/2 so that every value is
    exact in binary (Grid_ComputePressure.C:96-166)."""
    glob = dict(Gamma=1.5, DualEnergyFormalism=1, tiny_number=1e-20)
    cells = [(2.0, 10.0, 1.0, 2.0, 2.0, 3.0),    # DE: 0.5 * 2 * 3 = 3 (E and v unused)
             (2.0, 99.0, 5.0, 0.0, 0.0, 3.0),    # DE: still 3
             (4.0, 1.0, 0.0, 0.0, 0.0, 1e-30),   # 2e-30 below tiny_number -> 1e-20
             (4.0, 1.0, 0.0, 0.0, 0.0, -1.0)]    # negative -> 1e-20
    g = _cells(cells)
    p = oracle.compute_pressure(kind, g, glob)[0, 0, 3:3 + len(cells)]
    assert list(p) == [3.0, 3.0, 1e-20, 1e-20]
    # minimum support: max(p, K (gamma - 1) d^2) = max(3, 10 * 0.5 * 4) = 20 for the first cell
    p = oracle.compute_pressure(kind, g, glob, min_support=10.0)[0, 0, 3:3 + len(cells)]
    assert list(p) == [20.0, 20.0, 80.0, 80.0]
    # dual energy off: p = (gamma - 1) d (E - v^2/2) over the velocities of the
    # grid's rank: v = (1, 2, 2) -> v^2/2 = 0.5 (rank 1), 2.5 (rank 2), 4.5 (rank 3)
    glob0 = dict(glob, DualEnergyFormalism=0)
    for rank, want in ((1, 9.5), (2, 7.5), (3, 5.5)):
        g = _cells([(2.0, 10.0, 1.0, 2.0, 2.0, 0.0)], rank=rank, dual=False)
        p = oracle.compute_pressure(kind, g, glob0)
        assert p.reshape(-1)[np.ravel_multi_index((0, 0, 3), p.shape)] == want, rank
