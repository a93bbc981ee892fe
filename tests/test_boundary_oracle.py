"""Pins the oracle's restatement of grid::CopyZonesFromGrid
(oracle/boundary_oracle.py) by what the geometry fixes: on one grid spanning
a periodic domain the sibling copies are the periodic external boundary
fill, and on a domain cut into siblings every piece, ghosts included, equals
the window of the undivided grid after its periodic fill.  Pins the
restatement of ExternalBoundary::SetExternalBoundary (`external_boundary`):
a reflecting face is the periodic fill of the grid joined to its mirror
image with the normal velocity negated; outflow ghosts repeat the edge cell
and the fill is idempotent; inflow ghosts take the face's values; corners
come from the later axis; faces off the domain edge and periodic faces stay.
"""
import numpy as np
import pytest

import boundary_oracle as bo
from enzo_ppm import problems


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_single_grid_copies_are_the_periodic_fill(rank):
    n = [10, 8, 6][:rank]
    g, _, _ = problems.random_smooth(n, seed=rank, ncolours=2, rank=rank)
    a = g.copy()
    rng = np.random.default_rng(0)     # junk in the ghosts
    sl = a.active_slice()
    for f in a.fields:
        act = f[sl].copy()
        f[...] = rng.normal(size=f.shape)
        f[sl] = act
    b = a.copy()
    b.fill_periodic()
    start = [-a.ghost if d < rank else 0 for d in range(3)]
    bo.copy_zones([a], [start], a.active, periodic=True)
    for x, y in zip(a.fields, b.fields):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("parts", [(2, 2, 2), (3, 1, 2), (1, 4, 1)])
def test_siblings_equal_windows_of_the_whole(parts):
    g, _, _ = problems.turbulent(12, seed=4, ncolours=1)
    g.fill_periodic()
    pieces, starts = bo.split_domain(g, parts)
    bo.copy_zones(pieces, starts, g.active, periodic=True)
    gh = g.ghost
    for h, s in zip(pieces, starts):
        lo = [s[d] + gh for d in range(3)]       # g's index of the piece's cell 0
        w = tuple(slice(lo[d], lo[d] + h.dims[d]) for d in (2, 1, 0))
        for x, y in zip(h.fields, g.fields):
            assert np.array_equal(x, y[w])


def test_non_periodic_edges_stay():
    g, _, _ = problems.turbulent(12, seed=4, ncolours=0)
    g.fill_periodic()
    pieces, starts = bo.split_domain(g, (2, 1, 1))
    bo.copy_zones(pieces, starts, g.active, periodic=False)
    left = pieces[0]
    gh = g.ghost
    # the outer x ghosts of the left piece have no sibling: untouched (zero)
    assert np.all(left.fields[0][gh:-gh, gh:-gh, :gh] == 0)
    # its inner x ghosts come from the right piece's active cells
    assert np.array_equal(left.fields[0][gh:-gh, gh:-gh, -gh:],
                          pieces[1].fields[0][gh:-gh, gh:-gh, gh:2 * gh])


# ---------------------------------------------------------------------------
# external boundary (ExternalBoundary::SetExternalBoundary)

def _junk_ghosts(g, seed=0):
    rng = np.random.default_rng(seed)
    sl = g.active_slice()
    for f in g.fields:
        act = f[sl].copy()
        f[...] = rng.normal(size=f.shape)
        f[sl] = act
    return g


def _types(dim_types, rank=3):
    t = [bo.PERIODIC] * 6
    for d, (l, r) in dim_types.items():
        t[2 * d], t[2 * d + 1] = l, r
    return t


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_reflecting_face_is_the_periodic_fill_of_the_mirrored_grid(dim):
    from enzo_ppm.grid import Grid
    n = [6, 5, 7]
    g, _, _ = problems.random_smooth(n, seed=dim, ncolours=1)
    _junk_ghosts(g)
    m_act = list(n)
    m_act[dim] *= 2
    M = Grid(m_act, ncolours=g.ncolours, dual_energy=g.dual_energy)
    ax = 2 - dim
    vel = [g.Vel1Num, g.Vel2Num, g.Vel3Num][dim]
    for f, (fm, fg) in enumerate(zip(M.fields, g.fields)):
        act = fg[g.active_slice()]
        mirror = np.flip(act, axis=ax) * (-1.0 if f == vel else 1.0)
        fm[M.active_slice()] = np.concatenate([act, mirror], axis=ax)
    M.fill_periodic()
    bo.external_boundary(g, _types({dim: (bo.REFLECTING, bo.REFLECTING)}), [0, 0, 0], g.dims)
    gh = g.ghost
    # compare the layers along dim (whole extent there) inside the
    # transverse active window
    win_g = [slice(gh, gh + n[d]) for d in (2, 1, 0)]
    win_g[ax] = slice(0, g.dims[dim])
    win_m = list(win_g)
    win_m[ax] = slice(0, g.dims[dim])
    for fg, fm in zip(g.fields, M.fields):
        assert np.array_equal(fg[tuple(win_g)], fm[tuple(win_m)])


def test_outflow_repeats_the_edge_cell_and_is_idempotent():
    g, _, _ = problems.random_smooth([6, 5, 7], seed=3, ncolours=2)
    _junk_ghosts(g, 1)
    t = _types({0: (bo.OUTFLOW, bo.OUTFLOW), 1: (bo.OUTFLOW, bo.OUTFLOW),
                2: (bo.OUTFLOW, bo.OUTFLOW)})
    bo.external_boundary(g, t, [0, 0, 0], g.dims)
    gh, (nx, ny, nz) = g.ghost, g.active
    for f in g.fields:
        for i in range(gh):
            assert np.array_equal(f[:, :, i], f[:, :, gh])
            assert np.array_equal(f[:, :, gh + nx + i], f[:, :, gh + nx - 1])
            assert np.array_equal(f[:, i, :], f[:, gh, :])
            assert np.array_equal(f[i, :, :], f[gh, :, :])
            assert np.array_equal(f[gh + nz + i, :, :], f[gh + nz - 1, :, :])
    h = g.copy()
    bo.external_boundary(h, t, [0, 0, 0], g.dims)
    for x, y in zip(g.fields, h.fields):
        assert np.array_equal(x, y)


def test_inflow_faces_take_their_values_and_corners_come_from_the_later_axis():
    g, _, _ = problems.random_smooth([6, 5, 7], seed=4, ncolours=1)
    _junk_ghosts(g, 2)
    inflow = [[10.0 * f + k for k in range(6)] for f in range(g.nfields)]
    # x reflecting, y inflow: the (x ghost, y ghost) corners of Velocity1
    # take the y inflow value (x filled first, then y), not its negative
    t = _types({0: (bo.REFLECTING, bo.REFLECTING), 1: (bo.INFLOW, bo.INFLOW)})
    bo.external_boundary(g, t, [0, 0, 0], g.dims, inflow=inflow)
    gh, (nx, ny, nz) = g.ghost, g.active
    for f, arr in enumerate(g.fields):
        assert np.all(arr[:, :gh, :] == inflow[f][2])
        assert np.all(arr[:, gh + ny:, :] == inflow[f][3])
    assert np.all(g.fields[g.Vel1Num][:, :gh, :gh] == inflow[g.Vel1Num][2])


def test_faces_off_the_domain_edge_and_periodic_faces_stay():
    g, _, _ = problems.random_smooth([6, 5, 7], seed=5, ncolours=1)
    _junk_ghosts(g, 3)
    a = g.copy()
    # the grid sits 4 cells in from the domain's left x face and ends at its
    # right x face; y and z faces periodic
    bd = [g.dims[0] + 4, g.dims[1], g.dims[2]]
    bo.external_boundary(a, _types({0: (bo.OUTFLOW, bo.OUTFLOW)}), [4, 0, 0], bd)
    gh, nx = g.ghost, g.active[0]
    for x, y in zip(a.fields, g.fields):
        assert np.array_equal(x[:, :, :gh + nx], y[:, :, :gh + nx])      # left x ghosts kept
        assert np.array_equal(x[:, :, gh + nx:], np.repeat(y[:, :, gh + nx - 1:gh + nx], gh, 2))
