"""Flux-correction cases shared by the CPU-library and GPU tests: the
device's hipPPMCorrectForRefinedFluxes against the oracle's restatement,
every field bitwise, and the grid's own boundary fluxes."""
import numpy as np

import oracle
from enzo_ppm import device, problems
from enzo_ppm.grid import SubgridFluxes
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
