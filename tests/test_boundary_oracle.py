"""Pins the oracle's restatement of grid::CopyZonesFromGrid
(oracle/boundary_oracle.py) by what the geometry fixes: on one grid spanning
a periodic domain the sibling copies are the periodic external boundary
fill, and on a domain cut into siblings every piece, ghosts included, equals
the window of the undivided grid after its periodic fill."""
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
