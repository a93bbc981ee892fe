"""ORACLE -- test infrastructure only (tests/, smoke check, bench CPU baseline).

CPU restatement of grid::CopyZonesFromGrid (Grid_CopyZonesFromGrid.C:43-230,
the non-shearing branch) as SetBoundaryConditions uses it between the grids
of one level (EvolveLevel.C:785-791 -> SetBoundaryConditions.C), in integer
global cell indices instead of positions:

- a grid's extent including ghost zones is [start, start + GridDimension)
  per axis, `start` being the global index of its first (ghost) cell --
  the GridGlobalStart that grid::SolveHydroEquations hands to SolvePPM_DE
  (Grid_SolveHydroEquations.C:373-375);
- for every ordered pair (this, other), including this == other, and, on a
  periodic domain, every shift s in {-1, 0, 1} per axis of the domain size
  (the EdgeOffsets of Grid_CheckForOverlap.C), the overlap of this grid's
  whole extent shifted by s with the other grid's ACTIVE region is copied,
  every field, into this grid (copy3drel); this == other with no shift is
  skipped (it would copy the active region onto itself).

Copies read active cells and write ghost cells only (active regions of one
level do not overlap), so their order does not matter.

Pinned by tests/test_boundary_oracle.py: on one grid spanning a periodic
domain it equals the periodic external boundary fill
(ExternalBoundary::SetExternalBoundary, periodic case), and for a domain cut
into sibling grids every grid equals, ghosts included, the matching window
of the undivided grid after its periodic fill.

Also a restatement of ExternalBoundary::SetExternalBoundary
(ExternalBoundary_SetExternalBoundaryIO.C:50-1010, the branch built without
OOC_BOUNDARY) as grid::SetExternalBoundaryValues calls it for every baryon
field (Grid_SetExternalBoundaryValues.C:91-117), with one boundary type per
domain face (as ExternalBoundary::InitializeExternalBoundaryFace sets
them): `external_boundary`.  Pinned by tests/test_boundary_oracle.py: a
reflecting face equals the periodic fill of the grid joined to its mirror
image (normal velocity negated); outflow ghosts repeat the edge cell and the
fill is idempotent; inflow ghosts take the face's values; axes are filled in
order x, y, z (corners from the later axis); faces the grid does not share
with the domain, and periodic faces, are left alone.
"""
import itertools

import numpy as np


def copy_boxes(dims, starts, st, en, domain, periodic, rank):
    """Copy boxes (dst grid, src grid, dst start, src start, size), all in
    local (ghost-inclusive) indices, x first.  dims/starts/st/en: per grid,
    3 entries each (GridDimension, global start of cell 0, GridStartIndex,
    GridEndIndex); domain: active cells of the domain per axis."""
    G = len(dims)
    shifts = [(-1, 0, 1) if (periodic and d < rank) else (0,) for d in range(3)]
    boxes = []
    for a in range(G):
        for b in range(G):
            for s in itertools.product(*shifts):
                if a == b and s == (0, 0, 0):
                    continue
                lo, hi, ok = [0] * 3, [0] * 3, True
                for d in range(3):
                    if d >= rank:
                        lo[d], hi[d] = 0, 1
                        continue
                    A0 = starts[a][d] + s[d] * domain[d]
                    A1 = A0 + dims[a][d]
                    B0 = starts[b][d] + st[b][d]
                    B1 = starts[b][d] + en[b][d] + 1
                    lo[d], hi[d] = max(A0, B0), min(A1, B1)
                    if hi[d] <= lo[d]:
                        ok = False
                        break
                if not ok:
                    continue
                dst0 = [lo[d] - (starts[a][d] + s[d] * domain[d]) if d < rank else 0
                        for d in range(3)]
                src0 = [lo[d] - starts[b][d] if d < rank else 0 for d in range(3)]
                n = [hi[d] - lo[d] for d in range(3)]
                boxes.append((a, b, dst0, src0, n))
    return boxes


def copy_zones(grids, starts, domain, periodic=True):
    """Apply every copy box to the grids' fields (in place)."""
    rank = grids[0].rank
    dims = [g.dims for g in grids]
    st = [g.start for g in grids]
    en = [g.end for g in grids]
    for a, b, d0, s0, n in copy_boxes(dims, starts, st, en, domain, periodic, rank):
        dst = (slice(d0[2], d0[2] + n[2]), slice(d0[1], d0[1] + n[1]), slice(d0[0], d0[0] + n[0]))
        src = (slice(s0[2], s0[2] + n[2]), slice(s0[1], s0[1] + n[1]), slice(s0[0], s0[0] + n[0]))
        for fa, fb in zip(grids[a].fields, grids[b].fields):
            fa[dst] = fb[src]
    return grids


def split_domain(g, parts):
    """Cut grid g (active n per axis, periodic) into parts[0] x parts[1] x
    parts[2] sibling grids with the same ghost depth; returns (grids, global
    index of each piece's first cell, ghosts included).  The pieces' field
    ghosts are left zero; their acceleration (static) is g's window, so g's
    acceleration ghosts must be filled."""
    from enzo_ppm.grid import Grid
    gh = g.ghost
    act = g.active
    sizes = [act[d] // parts[d] for d in range(3)]
    out, starts = [], []
    for kz in range(parts[2]):
        for ky in range(parts[1]):
            for kx in range(parts[0]):
                k = (kx, ky, kz)
                h = Grid([sizes[d] for d in range(g.rank)], rank=g.rank, ghost=gh,
                         ncolours=g.ncolours, dual_energy=g.dual_energy, gravity=g.gravity,
                         dx=g.dx)
                lo = [k[d] * sizes[d] for d in range(3)]
                sl = tuple(slice(gh + lo[d], gh + lo[d] + sizes[d]) if d < g.rank else slice(0, 1)
                           for d in (2, 1, 0))
                asl = h.active_slice()
                for fh, fg in zip(h.fields, g.fields):
                    fh[asl] = fg[sl]
                if g.gravity:   # static: the whole window, ghosts included
                    wsl = tuple(slice(lo[d], lo[d] + sizes[d] + 2 * gh) if d < g.rank
                                else slice(0, 1) for d in (2, 1, 0))
                    for fh, fg in zip(h.accel, g.accel):
                        fh[...] = fg[wsl]
                out.append(h)
                starts.append([lo[d] - gh if d < g.rank else 0 for d in range(3)])
    return out, starts


# boundary types of the reference (typedefs.h: reflecting 0, outflow 1,
# inflow 2, periodic 3, shearing 4)
REFLECTING, OUTFLOW, INFLOW, PERIODIC, SHEARING = 0, 1, 2, 3, 4


def external_boundary(grid, types, grid_offset, boundary_dims, inflow=None, fields=None):
    """Ghost fill of the domain faces a grid touches, in place.

    types[2 * dim + face] (face 0 left, 1 right): the face's boundary type;
    grid_offset: nint((GridLeftEdge - DomainLeftEdge) / CellWidth) per axis;
    boundary_dims: the domain's cells per axis, ghosts included
    (BoundaryDimension); inflow[field][2 * dim + face]: inflow values.
    A left face is filled when grid_offset[dim] == 0, a right one when
    grid_offset[dim] + GridDimension[dim] == boundary_dims[dim]; the normal
    velocity (Velocity1/2/3 for x/y/z faces) changes sign at a reflecting
    face; periodic faces are left alone (the reference build leaves
    USE_PERIODIC undefined: periodic images come from CopyZonesFromGrid).
    `fields`: field numbers to fill (default all)."""
    vel = [grid.Vel1Num, grid.Vel2Num, grid.Vel3Num]
    fields = range(grid.nfields) if fields is None else fields
    S, E, GD = grid.start, grid.end, grid.dims
    for dim in range(grid.rank):
        if boundary_dims[dim] <= 1:
            continue
        for f in fields:
            a = np.moveaxis(grid.fields[f], 2 - dim, 0)   # sweep axis first
            sign = -1.0 if f == vel[dim] else 1.0
            n = E[dim] - S[dim] + 1
            if grid_offset[dim] == 0:            # left face (:145-260)
                t = types[2 * dim]
                for i in range(S[dim]):
                    if t == REFLECTING:
                        a[i] = sign * a[2 * S[dim] - 1 - i]
                    elif t == OUTFLOW:
                        a[i] = a[S[dim]]
                    elif t == INFLOW:
                        a[i] = inflow[f][2 * dim]
                    elif t == SHEARING:
                        a[i] = a[i + n]
                    elif t != PERIODIC:
                        raise ValueError(f"boundary type {t} not restated")
            if grid_offset[dim] + GD[dim] == boundary_dims[dim]:   # right face
                t = types[2 * dim + 1]
                for i in range(GD[dim] - E[dim] - 1):
                    x = E[dim] + 1 + i
                    if t == REFLECTING:
                        a[x] = sign * a[x - (2 * i + 1)]
                    elif t == OUTFLOW:
                        a[x] = a[x - 1 - i]
                    elif t == INFLOW:
                        a[x] = inflow[f][2 * dim + 1]
                    elif t == SHEARING:
                        a[x] = a[x - n]
                    elif t != PERIODIC:
                        raise ValueError(f"boundary type {t} not restated")
    return grid
