"""Pins the oracle's restatement of the AMR flux correction
(oracle/fluxcorr_oracle.c: grid::GetProjectedBoundaryFluxes and
grid::CorrectForRefinedFluxes for a proper subgrid) by what the algorithm
fixes: refined fluxes equal to the initial ones change nothing beyond the
conserved-form round trip; a flux difference moves exactly that much mass,
momentum and energy into the face cells; species keep their fractions;
total and gas energy are consistent afterwards; a correction that would
empty a cell is undone for every corrected field; faces on the grid's own
boundary leave the cells alone and replace the boundary fluxes; and the
projection conserves the fine fluxes and averages a uniform one."""
import numpy as np
import pytest

import oracle
from enzo_ppm import problems
from enzo_ppm.grid import SubgridFluxes

GS = [-3, -3, -3]          # global index of the grid's first cell: active starts at 0
BOX = ((5, 4, 6), (9, 10, 11))


def setup(seed=0, scale=1e-4, ncolours=2):
    g, glob, _ = problems.turbulent(16, seed=seed, ncolours=ncolours)
    ini = SubgridFluxes(g, [BOX])
    ref = SubgridFluxes(g, [BOX])
    rng = np.random.default_rng(seed)
    for key, (L, R) in ini.arrays[0].items():
        L[:] = rng.normal(0, 1e-3, L.shape)
        R[:] = rng.normal(0, 1e-3, R.shape)
        rL, rR = ref.arrays[0][key]
        rL[:] = L + rng.normal(0, scale, L.shape)
        rR[:] = R + rng.normal(0, scale, R.shape)
    return g, glob, ini, ref


def face_cells(g, dim):
    """(left cells, right cells) index tuples of the x/y/z faces of BOX."""
    s = [BOX[0][d] - GS[d] for d in range(3)]
    e = [BOX[1][d] - GS[d] for d in range(3)]
    lo, hi = list(s), list(e)
    lo[dim] = hi[dim] = s[dim] - 1
    left = tuple(slice(lo[d], hi[d] + 1) for d in (2, 1, 0))
    lo[dim] = hi[dim] = e[dim] + 1
    right = tuple(slice(lo[d], hi[d] + 1) for d in (2, 1, 0))
    return left, right


def face(arr, g, dim):
    """A face flux array as (z, y, x) block."""
    n = [BOX[1][d] - BOX[0][d] + 1 for d in range(3)]
    n[dim] = 1
    return arr.reshape(n[2], n[1], n[0])


def test_equal_fluxes_change_nothing_but_rounding():
    g, glob, ini, _ = setup()
    same = SubgridFluxes(g, [BOX])
    for key, (L, R) in ini.arrays[0].items():
        same.arrays[0][key][0][:] = L
        same.arrays[0][key][1][:] = R
    a = g.copy()
    assert oracle.correct_for_refined_fluxes(a, GS, ini.structs[0], same.structs[0], None,
                                             glob, [0, 1]) == 1
    assert np.array_equal(a.fields[g.DensNum], g.fields[g.DensNum])
    for k in (g.Vel1Num, g.Vel2Num, g.Vel3Num, g.GENum, g.colnum[0]):
        rel = np.abs(a.fields[k] - g.fields[k]) / np.maximum(np.abs(g.fields[k]), 1e-300)
        assert rel.max() <= 4.5e-16, (k, rel.max())
    assert np.array_equal(a.fields[g.colnum[1]], g.fields[g.colnum[1]])   # passive


def test_differences_move_conserved_quantities():
    g, glob, ini, ref = setup(seed=1)
    a = g.copy()
    assert oracle.correct_for_refined_fluxes(a, GS, ini.structs[0], ref.structs[0], None,
                                             glob, [0, 1]) == 1
    D, E, GE = g.DensNum, g.TENum, g.GENum
    for dim in range(3):
        left, right = face_cells(g, dim)
        iL = lambda f: face(ini.arrays[0][(f, dim)][0], g, dim)  # noqa: E731
        iR = lambda f: face(ini.arrays[0][(f, dim)][1], g, dim)  # noqa: E731
        rL = lambda f: face(ref.arrays[0][(f, dim)][0], g, dim)  # noqa: E731
        rR = lambda f: face(ref.arrays[0][(f, dim)][1], g, dim)  # noqa: E731
        # density: + (initial - refined) on the left, - on the right
        assert np.allclose(a.fields[D][left] - g.fields[D][left], iL(D) - rL(D), rtol=0,
                           atol=1e-15)
        assert np.allclose(a.fields[D][right] - g.fields[D][right], -(iR(D) - rR(D)), rtol=0,
                           atol=1e-15)
        for cells, sgn, i_, r_ in ((left, 1, iL, rL), (right, -1, iR, rR)):
            d0, d1 = g.fields[D][cells], a.fields[D][cells]
            # momentum rho v moves by the flux difference
            for v in (g.Vel1Num, g.Vel2Num, g.Vel3Num):
                mom = d1 * a.fields[v][cells] - d0 * g.fields[v][cells]
                assert np.allclose(mom, sgn * (i_(v) - r_(v)), rtol=1e-10, atol=1e-15)
            # species keep their fraction of the density
            c = g.colnum[0]
            assert np.allclose(a.fields[c][cells] / d1, g.fields[c][cells] / d0, rtol=1e-14,
                               atol=0)
            # gas energy (rho e) moves by its flux difference; total energy is
            # made consistent with it
            ge = d1 * a.fields[GE][cells] - d0 * g.fields[GE][cells]
            assert np.allclose(ge, sgn * (i_(GE) - r_(GE)), rtol=1e-10, atol=1e-15)
            ke = 0.5 * sum(a.fields[v][cells] ** 2 for v in (g.Vel1Num, g.Vel2Num, g.Vel3Num))
            assert np.allclose(a.fields[E][cells], a.fields[GE][cells] + ke, rtol=1e-15,
                               atol=0)


def test_correction_emptying_a_cell_is_undone():
    g, glob, ini, ref = setup(seed=2, scale=0.0)
    D = g.DensNum
    dim = 1
    left, _ = face_cells(g, dim)
    # make the refined density flux so small on the left face that the first
    # face cell would go negative: the correction there is undone
    L = ref.arrays[0][(D, dim)][0]
    Lface = face(L, g, dim)
    cell = tuple(s.start for s in left)          # (z, y, x) of the face's first cell
    Lface[0, 0, 0] = face(ini.arrays[0][(D, dim)][0], g, dim)[0, 0, 0] + 2 * g.fields[D][cell]
    for v in (g.Vel1Num, g.TENum):
        face(ref.arrays[0][(v, dim)][0], g, dim)[0, 0, 0] += 1.0
    a = g.copy()
    assert oracle.correct_for_refined_fluxes(a, GS, ini.structs[0], ref.structs[0], None,
                                             glob, [0, 1]) == 1
    assert a.fields[D][cell] == g.fields[D][cell]
    for v in (g.Vel1Num, g.GENum):
        assert abs(a.fields[v][cell] - g.fields[v][cell]) <= 4.5e-16 * abs(g.fields[v][cell])
    # and the refined fluxes of every corrected field were reset there
    for f in (D, g.TENum, g.GENum, g.Vel1Num):
        assert face(ref.arrays[0][(f, dim)][0], g, dim)[0, 0, 0] == \
            face(ini.arrays[0][(f, dim)][0], g, dim)[0, 0, 0]


def test_faces_on_the_grid_boundary_replace_boundary_fluxes():
    g, glob, _ = problems.turbulent(16, seed=3, ncolours=0)
    box = ((0, 4, 6), (9, 10, 11))                 # touches the active x = 0 face
    ini, ref = SubgridFluxes(g, [box]), SubgridFluxes(g, [box])
    own = SubgridFluxes(g, [((0, 0, 0), (15, 15, 15))])   # BoundaryFluxesThisTimeStep
    rng = np.random.default_rng(3)
    for key, (L, R) in ini.arrays[0].items():
        L[:] = rng.normal(0, 1e-3, L.shape)
        R[:] = rng.normal(0, 1e-3, R.shape)
        ref.arrays[0][key][0][:] = L + 1e-4
        ref.arrays[0][key][1][:] = R + 1e-4
    a = g.copy()
    assert oracle.correct_for_refined_fluxes(a, GS, ini.structs[0], ref.structs[0],
                                             own.structs[0], glob, [], 1) == 1
    # the x = -1 ghost layer is not corrected; the grid's own left x fluxes
    # take the refined ones over the face
    assert np.array_equal(a.fields[g.DensNum][6 + 3:12 + 3, 4 + 3:11 + 3, 2],
                          g.fields[g.DensNum][6 + 3:12 + 3, 4 + 3:11 + 3, 2])
    Lown = own.arrays[0][(g.DensNum, 0)][0].reshape(16, 16, 1)
    assert np.array_equal(Lown[6:12, 4:11, 0],
                          ref.arrays[0][(g.DensNum, 0)][0].reshape(6, 7, 1)[:, :, 0])


@pytest.mark.parametrize("r", [2, 4])
def test_projection_averages_and_conserves(r):
    parent, _, _ = problems.turbulent(16, seed=4, ncolours=0)
    cbox = ((2, 4, 6), (5, 7, 9))
    fine_box = (tuple(c * r for c in cbox[0]), tuple((c + 1) * r - 1 for c in cbox[1]))
    fine_grid, _, _ = problems.turbulent(8, seed=4, ncolours=0)
    fine = SubgridFluxes(fine_grid, [fine_box])
    proj = SubgridFluxes(parent, [cbox])
    rng = np.random.default_rng(r)
    for (f, dim), (L, R) in fine.arrays[0].items():
        if f == 0:
            L[:] = 3.0                       # uniform: averages to 3 / r
            R[:] = rng.normal(0, 1.0, R.shape)
        else:
            L[:] = rng.normal(0, 1.0, L.shape)
            R[:] = rng.normal(0, 1.0, R.shape)
    oracle.project_fluxes(fine.structs[0], 3, parent.nfields, [r, r, r], proj.structs[0])
    for (f, dim), (L, R) in proj.arrays[0].items():
        fL, fR = fine.arrays[0][(f, dim)]
        if f == 0:
            assert np.allclose(L, 3.0 / r, rtol=1e-15, atol=0)
        assert np.isclose(R.sum() * r ** 3, fR.sum(), rtol=1e-12, atol=1e-12)
        assert np.isclose(L.sum() * r ** 3, fL.sum(), rtol=1e-12, atol=1e-12)
    P = proj.structs[0]
    assert [P.LeftFluxStartGlobalIndex[0][d] for d in range(3)] == list(cbox[0])
