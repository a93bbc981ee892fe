"""Flux-correction cases shared by the CPU-library and GPU tests: the
device's hipPPMCorrectForRefinedFluxes against the oracle's restatement,
every field bitwise, and the grid's own boundary fluxes."""
import numpy as np

import oracle
from enzo_ppm import device, problems
from enzo_ppm.grid import Grid, SubgridFluxes
from helpers import field_report

GS = [-3, -3, -3]

CASES = {
    # name: (box, perturbation scale, emptying cells, correct parent boundary)
    "interior": (((5, 4, 6), (9, 10, 11)), 1e-4, False, 0),
    "undo": (((3, 3, 3), (8, 12, 9)), 1e-4, True, 0),
    "on_left_x_edge": (((0, 4, 6), (9, 10, 11)), 1e-4, False, 1),
    "on_right_z_edge": (((2, 2, 4), (13, 13, 15)), 1e-4, False, 1),
}


def case(name, n=16, seed=0):
    box, scale, empty, cpbf = CASES[name]
    g, glob, _ = problems.turbulent(n, seed=seed, ncolours=3)
    ini, ref = SubgridFluxes(g, [box]), SubgridFluxes(g, [box])
    own = SubgridFluxes(g, [((0, 0, 0), (n - 1, n - 1, n - 1))])
    rng = np.random.default_rng(seed)
    for key, (L, R) in ini.arrays[0].items():
        L[:] = rng.normal(0, 1e-3, L.shape)
        R[:] = rng.normal(0, 1e-3, R.shape)
        ref.arrays[0][key][0][:] = L + rng.normal(0, scale, L.shape)
        ref.arrays[0][key][1][:] = R + rng.normal(0, scale, R.shape)
    if empty:  # some refined density fluxes empty their cells: undone there
        for dim in range(3):
            L = ref.arrays[0][(g.DensNum, dim)][0]
            R = ref.arrays[0][(g.DensNum, dim)][1]
            L[::5] += 50.0
            R[1::7] -= 50.0
    return g, glob, ini, ref, own, cpbf


def run_case(lib, name, kinds=(0, 1, 2)):
    g, glob, ini, ref, own, cpbf = case(name)
    a, b = g.copy(), g.copy()
    own_b = SubgridFluxes(g, [((0, 0, 0), (15, 15, 15))])
    # the oracle modifies the refined fluxes it undoes: give it its own copy
    ref_o = SubgridFluxes(g, [CASES[name][0]])
    for key, (L, R) in ref.arrays[0].items():
        ref_o.arrays[0][key][0][:] = L
        ref_o.arrays[0][key][1][:] = R
    st = oracle.correct_for_refined_fluxes(b, GS, ini.structs[0], ref_o.structs[0],
                                           own_b.structs[0], glob, list(kinds), cpbf)
    assert st == 1
    dg = device.DeviceGrid(a, glob, lib=lib, global_start=GS)
    try:
        dg.correct_for_refined_fluxes(ini.structs[0], ref.structs[0], own.structs[0], kinds, cpbf)
        assert dg.download() == 1
    finally:
        dg.close()
    rep = field_report(a, b)
    assert all(r[1] for r in rep), rep
    assert np.array_equal(own.flat(), own_b.flat())
    changed = sum(int(np.count_nonzero(x != y)) for x, y in zip(a.fields, g.fields))
    return changed


def run_level_case(lib, name, steps=0, kinds=(0, 1, 2)):
    """The same correction on grid 1 of a two-grid batched level
    (hipPPMLevelCorrectForRefinedFluxes), after `steps` level steps of a
    fixed dt (both buffer parities): grid 1 equals the oracle's step(s) and
    correction bitwise, grid 0 the oracle's step(s) alone."""
    from enzo_ppm.level import Level
    from helpers import level_reference
    g, glob, ini, ref, own, cpbf = case(name)
    other, _, _ = problems.turbulent(12, seed=3, ncolours=3)
    grids = [other, g]
    for x in grids:
        x.fill_periodic()
    b = [x.copy() for x in grids]
    if steps:
        level_reference(b, glob, steps, dt=1e-3)
    own_b = SubgridFluxes(g, [((0, 0, 0), (15, 15, 15))])
    ref_o = SubgridFluxes(g, [CASES[name][0]])
    for key, (L, R) in ref.arrays[0].items():
        ref_o.arrays[0][key][0][:] = L
        ref_o.arrays[0][key][1][:] = R
    st = oracle.correct_for_refined_fluxes(b[1], GS, ini.structs[0], ref_o.structs[0],
                                           own_b.structs[0], glob, list(kinds), cpbf)
    assert st == 1
    lv = Level([x.copy() for x in grids], glob, lib=lib, global_starts=[(0, 0, 0), GS])
    try:
        for s in range(steps):
            lv.step(s, 1e-3, 0.4)
        lv.correct_for_refined_fluxes(1, ini.structs[0], ref.structs[0], own.structs[0], kinds, cpbf)
        assert lv.status()[0] == 1
        lv.download()
        for x, y in zip(lv.grids, b):
            rep = field_report(x, y)
            assert all(r[1] for r in rep), rep
    finally:
        lv.close()
    assert np.array_equal(own.flat(), own_b.flat())


def run_projection(lib, r):
    """hipPPMProjectBoundaryFluxes against the oracle, bitwise."""
    parent, _, _ = problems.turbulent(16, seed=4, ncolours=1)
    cbox = ((2, 4, 6), (5, 7, 9))
    fine_box = (tuple(c * r for c in cbox[0]), tuple((c + 1) * r - 1 for c in cbox[1]))
    fine = SubgridFluxes(parent, [fine_box])
    a, b = SubgridFluxes(parent, [cbox]), SubgridFluxes(parent, [cbox])
    rng = np.random.default_rng(r)
    for (f, dim), (L, R) in fine.arrays[0].items():
        L[:] = rng.normal(0, 1.0, L.shape)
        R[:] = rng.normal(0, 1.0, R.shape)
    oracle.project_fluxes(fine.structs[0], 3, parent.nfields, [r, r, r], b.structs[0])
    import ctypes as C
    from enzo_ppm import abi
    st = lib.hipPPMProjectBoundaryFluxes(C.byref(fine.structs[0]), 3, parent.nfields,
                                         abi.iarr([r, r, r]), C.byref(a.structs[0]))
    assert st == 1
    assert np.array_equal(a.flat(), b.flat())
    for d in range(3):
        for j in range(3):
            assert a.structs[0].LeftFluxStartGlobalIndex[d][j] == \
                b.structs[0].LeftFluxStartGlobalIndex[d][j]


# ---------------------------------------------------------------------------
# SUBling correction (UpdateFromFinerGrids.C:180-198): two neighbouring
# coarse grids A (global x 0..15) and B (x 16..31) of 16^3 cells, cut out
# of one 32 x 16 x 16 grid AB, and a finer grid G (its box in coarse cells)
# that is a child of one of them and touches the other's face.  The grid
# being corrected is B for side "left" (G, a child of A, touches B's left
# face) and A for side "right" (G, a child of B, touches A's right face).
AB_N = (32, 16, 16)
GBOX = {"left": ((10, 4, 6), (15, 10, 11)), "right": ((16, 4, 6), (20, 10, 11))}
TARGET_X = {"left": 16, "right": 0}      # global x of the corrected grid's first active cell


def x_window(g, x0, nx):
    """Grid of nx active cells along x holding g's cells from g's local x
    index x0 on (ghosts included)."""
    h = Grid([nx, g.active[1], g.active[2]], ncolours=g.ncolours, dual_energy=g.dual_energy,
             dx=g.dx)
    for fh, fg in zip(h.fields, g.fields):
        fh[...] = fg[:, :, x0:x0 + nx + 2 * g.ghost]
    return h


def copy_fluxes(src, grid, box):
    out = SubgridFluxes(grid, [box])
    for key, (L, R) in src.arrays[0].items():
        out.arrays[0][key][0][:] = L
        out.arrays[0][key][1][:] = R
    return out


def subling_setup(side, seed=0, empty=None):
    """(AB, glob, ini_AB, ref): AB's proper-subgrid fluxes of G (initial and
    projected refined ones); `empty` = field whose refined flux empties the
    first face cell of the corrected grid (its correction is undone)."""
    AB, glob, _ = problems.random_smooth(list(AB_N), seed=seed, ncolours=3)
    box = GBOX[side]
    ini, ref = SubgridFluxes(AB, [box]), SubgridFluxes(AB, [box])
    rng = np.random.default_rng(seed)
    for key, (L, R) in ini.arrays[0].items():
        L[:] = rng.normal(0, 1e-3, L.shape)
        R[:] = rng.normal(0, 1e-3, R.shape)
        ref.arrays[0][key][0][:] = L + rng.normal(0, 1e-4, L.shape)
        ref.arrays[0][key][1][:] = R + rng.normal(0, 1e-4, R.shape)
    if empty is not None:
        # the face of G next to the corrected grid's cell at (x, y 4, z 6):
        # G's right x face for "left" (cell x 16: += -(ini - ref)), its left
        # x face for "right" (cell x 15: += ini - ref); face position 0
        x = 16 if side == "left" else 15
        cell = (6 + 3, 4 + 3, x + 3)
        q = AB.fields[empty][cell] * (1.0 if empty == AB.DensNum else AB.fields[AB.DensNum][cell])
        if side == "left":
            ref.arrays[0][(empty, 0)][1][0] = ini.arrays[0][(empty, 0)][1][0] - 2.0 * abs(q)
        else:
            ref.arrays[0][(empty, 0)][0][0] = ini.arrays[0][(empty, 0)][0][0] + 2.0 * abs(q)
    return AB, glob, ini, ref


def subling_target(AB, side, ini_AB, ref, wrap=False):
    """The corrected grid as a SUBling call's inputs: (grid, GridGlobalStart,
    own boundary fluxes, refined fluxes in the level's index space,
    GlobalDim, PeriodicFace, G's box).  Its own fluxes on the face G touches equal
    AB's initial fluxes of G there, so the SUBling correction must equal
    AB's proper-subgrid correction on those cells.  wrap: the same two grids
    placed across the periodic boundary of a 32-cell domain (the corrected
    grid at one end, G at the other)."""
    x0 = TARGET_X[side]
    T = x_window(AB, x0, 16)
    gdim = [32, 16, 16]
    if not wrap:
        tx = x0
        gbox = GBOX[side]
    elif side == "left":            # B at x 0..15, G at the domain's right end
        tx, gbox = 0, ((26, 4, 6), (31, 10, 11))
    else:                           # A at x 16..31, G at the domain's left end
        tx, gbox = 16, ((0, 4, 6), (4, 10, 11))
    gs = [tx - 3, -3, -3]
    own = SubgridFluxes(T, [((tx, 0, 0), (tx + 15, 15, 15))])
    rng = np.random.default_rng(99)
    for key, (L, R) in own.arrays[0].items():
        L[:] = rng.normal(0, 1e-3, L.shape)
        R[:] = rng.normal(0, 1e-3, R.shape)
    # the touched face: B's left (own left x flux) = AB's right x flux of G,
    # A's right (own right x flux) = AB's left x flux of G
    box = GBOX[side]
    ny, nz = box[1][1] - box[0][1] + 1, box[1][2] - box[0][2] + 1
    for f in range(AB.nfields):
        src = ini_AB.arrays[0][(f, 0)][1 if side == "left" else 0].reshape(nz, ny)
        dst = own.arrays[0][(f, 0)][0 if side == "left" else 1].reshape(16, 16)
        dst[box[0][2]:box[1][2] + 1, box[0][1]:box[1][1] + 1] = src
    refT = copy_fluxes(ref, T, gbox)
    return T, gs, own, refT, gdim, [1] * 6 if wrap else [0] * 6, gbox


KINDS = (0, 1, 2)


def oracle_subling(T, gs, own, refT, gdim, periodic, gbox, glob, kinds=KINDS):
    """The oracle's SUBling correction of a copy of T (and of the refined
    fluxes, which it resets where it undoes)."""
    t = T.copy()
    r = copy_fluxes(refT, T, gbox)
    st = oracle.correct_for_refined_fluxes(t, gs, own.structs[0], r.structs[0], None, glob,
                                           list(kinds), 0, 1, gdim, periodic)
    assert st == 1
    return t, r


SUBLING_CASES = {
    # name: (side, emptied field or None, across the periodic boundary)
    "left": ("left", None, False),
    "right": ("right", None, False),
    "left_undo_density": ("left", 0, False),
    "right_undo_energy": ("right", 1, False),
    "left_periodic": ("left", None, True),
    "right_periodic": ("right", 1, True),
}


def run_subling_case(lib, name):
    """hipPPMCorrectForRefinedFluxes with SUBlingGrid against the oracle,
    every field bitwise."""
    side, empty, wrap = SUBLING_CASES[name]
    AB, glob, ini, ref = subling_setup(side, seed=3, empty=empty)
    T, gs, own, refT, gdim, per, gbox = subling_target(AB, side, ini, ref, wrap)
    b, _ = oracle_subling(T, gs, own, refT, gdim, per, gbox, glob)
    a = T.copy()
    dg = device.DeviceGrid(a, glob, lib=lib, global_start=gs)
    try:
        dg.correct_for_refined_fluxes(own.structs[0], refT.structs[0], None, KINDS, 0,
                                      subling=1, global_dim=gdim, periodic_face=per)
        assert dg.download() == 1
    finally:
        dg.close()
    rep = field_report(a, b)
    assert all(r[1] for r in rep), rep
    return sum(int(np.count_nonzero(x != y)) for x, y in zip(a.fields, T.fields))
