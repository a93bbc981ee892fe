"""Pins the oracle's restatement of the AMR flux correction
(oracle/fluxcorr_oracle.c: grid::GetProjectedBoundaryFluxes and
grid::CorrectForRefinedFluxes for a proper subgrid) by what the algorithm
fixes: refined fluxes equal to the initial ones change nothing beyond the
conserved-form round trip; a flux difference moves exactly that much mass,
momentum and energy into the face cells; species keep their fractions;
total and gas energy are consistent afterwards; a correction that would
empty a cell is undone for every corrected field; faces on the grid's own
boundary leave the cells alone and replace the boundary fluxes; and the
projection conserves the fine fluxes and averages a uniform one.  For a
SUBling (SUBlingGrid TRUE): the correction of a grid at the face a finer
grid of its neighbour touches equals, bit for bit, the proper-subgrid
correction of one grid holding both neighbours on the cells next to that
face (undo included), a periodic wrap of the same placement gives the same
result, grids that do not touch change nothing, and the reference's undo on
a SUBling's right-hand face subtracts the amount of the field it found
non-positive from the fields corrected before it."""
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


# ---------------------------------------------------------------------------
# SUBlings (SUBlingGrid TRUE): the correction of a grid at a face that a
# finer grid of its NEIGHBOUR touches must equal, cell for cell, the
# proper-subgrid correction of one grid holding both neighbours, on the
# cells next to that face; a periodic wrap must not change it.
import fluxcorr_checks as fc  # noqa: E402


def proper_joined(AB, glob, side, ini, ref):
    a = AB.copy()
    r = fc.copy_fluxes(ref, AB, fc.GBOX[side])
    assert oracle.correct_for_refined_fluxes(a, GS, ini.structs[0], r.structs[0], None, glob,
                                             list(fc.KINDS)) == 1
    return a, r


def layers(side):
    """(x of the corrected layer in the target, in AB; x of the target's
    other face layer), local indices; y, z window of G's face."""
    yz = (slice(4 + 3, 11 + 3), slice(6 + 3, 12 + 3))
    return (3, 19, 18, yz) if side == "left" else (18, 18, 3, yz)


@pytest.mark.parametrize("side,empty", [("left", None), ("right", None), ("left", 0),
                                        ("right", 0), ("left", 1)])
def test_subling_equals_proper_correction_of_joined_grid(side, empty):
    AB, glob, ini, ref = fc.subling_setup(side, seed=5, empty=empty)
    a, ra = proper_joined(AB, glob, side, ini, ref)
    T, gs, own, refT, gdim, per, gbox = fc.subling_target(AB, side, ini, ref)
    t, rt = fc.oracle_subling(T, gs, own, refT, gdim, per, gbox, glob)
    xt, xab, xo, (ys, zs) = layers(side)
    for f in range(AB.nfields):
        assert np.array_equal(t.fields[f][zs, ys, xt], a.fields[f][zs, ys, xab]), f
        if empty is None:
            assert not np.array_equal(t.fields[f][zs, ys, xt], T.fields[f][zs, ys, xt]) or \
                f in (AB.colnum[1],), f
        # the other face layer of the target only makes the conserved-form
        # round trip; every other cell is untouched
        rel = np.abs(t.fields[f][zs, ys, xo] - T.fields[f][zs, ys, xo]) / \
            np.maximum(np.abs(T.fields[f][zs, ys, xo]), 1e-300)
        assert rel.max() <= 1e-15, (f, rel.max())
        m = np.ones(T.shape, bool)
        m[zs, ys, xt] = False
        m[zs, ys, xo] = False
        assert np.array_equal(t.fields[f][m], T.fields[f][m]), f
    # where the correction was undone, the refined fluxes were reset alike
    k = 1 if side == "left" else 0
    for f in range(AB.nfields):
        assert np.array_equal(rt.arrays[0][(f, 0)][k], ra.arrays[0][(f, 0)][k]), f
    if empty is not None:   # the undo happened: the emptying flux was reset
        assert rt.arrays[0][(empty, 0)][k][0] == ini.arrays[0][(empty, 0)][k][0]
        assert refT.arrays[0][(empty, 0)][k][0] != ini.arrays[0][(empty, 0)][k][0]


AB_TE = 1      # field number of the total energy in these grids


def test_subling_right_face_energy_undo_keeps_reference_amount():
    """On a SUBling's right-hand face the reference's undo subtracts the
    amount of the field found non-positive from every field corrected before
    it (Grid_CorrectForRefinedFluxes.C:638-652): when the total energy is
    undone, the density loses the energy's amount, not its own."""
    side = "right"
    AB, glob, ini, ref = fc.subling_setup(side, seed=5, empty=AB_TE)
    T, gs, own, refT, gdim, per, gbox = fc.subling_target(AB, side, ini, ref)
    t, _ = fc.oracle_subling(T, gs, own, refT, gdim, per, gbox, glob)
    a, _ = proper_joined(AB, glob, side, ini, ref)
    cell = (6 + 3, 4 + 3, 18)
    D, E = AB.DensNum, AB.TENum
    IR = {f: own.arrays[0][(f, 0)][1].reshape(16, 16)[6, 4] for f in (D, E)}
    RL = {f: refT.arrays[0][(f, 0)][0][0] for f in (D, E)}
    aD, aE = IR[D] - RL[D], IR[E] - RL[E]
    r0 = T.fields[D][cell]
    assert T.fields[E][cell] * r0 + aE <= 0          # the energy correction was undone
    assert t.fields[D][cell] == (r0 + aD) - aE
    assert a.fields[D][cell] == (r0 + aD) - aD       # a proper subgrid's undo
    assert t.fields[D][cell] != a.fields[D][cell]


@pytest.mark.parametrize("side", ["left", "right"])
def test_subling_across_periodic_boundary(side):
    AB, glob, ini, ref = fc.subling_setup(side, seed=7, empty=None)
    T, gs, own, refT, gdim, per, gbox = fc.subling_target(AB, side, ini, ref, wrap=False)
    t0, _ = fc.oracle_subling(T, gs, own, refT, gdim, per, gbox, glob)
    T, gs, own, refT, gdim, per, gbox = fc.subling_target(AB, side, ini, ref, wrap=True)
    t1, _ = fc.oracle_subling(T, gs, own, refT, gdim, per, gbox, glob)
    for f in range(AB.nfields):
        assert np.array_equal(t0.fields[f], t1.fields[f]), f
    assert any(not np.array_equal(x, y) for x, y in zip(t1.fields, T.fields))
    # the same placement on a non-periodic domain: the grids do not touch
    t2, _ = fc.oracle_subling(T, gs, own, refT, gdim, [0] * 6, gbox, glob)
    for f in range(AB.nfields):
        assert np.array_equal(t2.fields[f], T.fields[f]), f


def test_subling_not_touching_changes_nothing():
    AB, glob, ini, ref = fc.subling_setup("left", seed=2)
    T, gs, own, refT, gdim, per, _ = fc.subling_target(AB, "left", ini, ref)
    far = ((3, 4, 6), (8, 10, 11))            # G two cells away from B's face
    refF = fc.copy_fluxes(refT, T, far)
    t, _ = fc.oracle_subling(T, gs, own, refF, gdim, per, far, glob)
    for f in range(AB.nfields):
        assert np.array_equal(t.fields[f], T.fields[f]), f
