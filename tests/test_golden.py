"""The oracle against golden fixtures written by the reference Fortran
(tests/golden/make_golden.py): bitwise equality of every baryon field after
each step, of the Courant time steps and of the captured face fluxes."""
import ctypes as C
import glob as globmod
import os

import numpy as np
import pytest

import oracle
from enzo_ppm.grid import Grid, SubgridFluxes
from helpers import bit_equal, fallback_differences

GOLD = os.path.dirname(os.path.abspath(__file__)) + "/golden"
STEP_FILES = sorted(globmod.glob(os.path.join(GOLD, "step_*.npz")))


def load_case(path):
    z = np.load(path, allow_pickle=False)
    glob = {str(k): (int(v) if str(k) != "Gamma" else float(v))
            for k, v in zip(z["glob_keys"], z["glob_vals"])}
    g = Grid(list(z["active"]), rank=int(z["rank"]), ncolours=int(z["ncolours"]),
             dual_energy=bool(z["dual"]), gravity=bool(z["gravity"]), dx=list(z["dx"]))
    for k in range(g.nfields):
        g.fields[k][...] = z[f"in0_f{k}"]
    if g.gravity:
        for a in range(3):
            g.accel[a][...] = z[f"accel{a}"]
    return z, g, glob


def replay(z, g, glob, step_fn, name):
    """Replays the fixture's steps with step_fn(grid, cycle, fluxes)."""
    periodic = "blast" not in name
    boxes = [tuple(map(tuple, b)) for b in z["boxes"]] if "boxes" in z.files else None
    for s in range(int(z["nsteps"])):
        if periodic:
            g.fill_periodic()
        g.dt = oracle.compute_time_step("port", g, glob, float(z["cfl"]))
        assert g.dt == float(z[f"dt{s}"]), (s, g.dt, float(z[f"dt{s}"]))
        fl = SubgridFluxes(g, boxes) if boxes else None
        step_fn(g, s, fl)
        for k in range(g.nfields):
            assert bit_equal(g.fields[k], z[f"out{s}_f{k}"]), (name, s, k)
        if fl is not None:
            assert bit_equal(fl.flat(), z[f"flux{s}"]), (name, s)


@pytest.mark.parametrize("path", STEP_FILES, ids=[os.path.basename(p) for p in STEP_FILES])
def test_oracle_matches_golden(path):
    z, g, glob = load_case(path)

    def step(g, s, fl):
        st, fb = oracle.solve_ppm_de("port", g, glob, cycle=s, fluxes=fl)
        assert st == 1 and fb == 0

    replay(z, g, glob, step, os.path.basename(path))


def test_twoshock_golden():
    z = np.load(os.path.join(GOLD, "twoshock_512.npz"), allow_pickle=False)
    idim = z["dls"].shape[1]
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    i = lambda v: C.byref(C.c_int64(v))
    d = lambda v: C.byref(C.c_double(v))
    pl, pr = z["pls"].copy(), z["prs"].copy()
    pbar, ubar = np.zeros((1, idim)), np.zeros((1, idim))
    dls, drs, uls, urs = (np.ascontiguousarray(z[k]) for k in ("dls", "drs", "uls", "urs"))
    oracle.lib("port").orc_twoshock(p(dls), p(drs), p(pl), p(pr), p(uls), p(urs), i(idim), i(1),
                                    i(1), i(idim), i(1), i(1), d(0.0), d(5 / 3), d(1e-20), i(0),
                                    p(pbar), p(ubar), i(0), p(pbar), i(1), d(1e-3))
    assert bit_equal(pbar, z["pbar"]) and bit_equal(ubar, z["ubar"])


FB_FILES = sorted(globmod.glob(os.path.join(GOLD, "fallback_*.npz")))


@pytest.mark.parametrize("path", FB_FILES, ids=[os.path.basename(p) for p in FB_FILES])
def test_oracle_matches_golden_fallback(path):
    """Per-cell HLL fallback steps written by the reference: same status and
    fallback count, same fields except the gas energy of the cells after
    the fallback cells (read from unset locals by the reference)."""
    z = np.load(path, allow_pickle=False)
    glob = {str(k): (int(v) if str(k) != "Gamma" else float(v))
            for k, v in zip(z["glob_keys"], z["glob_vals"])}
    g = Grid(list(z["active"]), rank=int(z["rank"]), ncolours=int(z["ncolours"]), dual_energy=True,
             dx=list(z["dx"]))
    for d in range(3):
        g.cellwidth[d][...] = z[f"cellwidth{d}"]
    for k in range(g.nfields):
        g.fields[k][...] = z[f"in_f{k}"]
    g.dt = float(z["dt"])
    st, fb = oracle.solve_ppm_de("port", g, glob, cycle=int(z["cycle"]))
    assert (st, fb) == (int(z["status"]), int(z["fallbacks"]))
    fallback_differences(g, [z[f"out_f{k}"] for k in range(g.nfields)], fb)


FLATTEN_FILES = [p for p in STEP_FILES if "flatten" in os.path.basename(p)]


@pytest.mark.parametrize("path", FLATTEN_FILES, ids=[os.path.basename(p) for p in FLATTEN_FILES])
def test_flattening_fixtures_exercise_their_type(path):
    """The flattening fixtures pin their type: the same steps with
    PPMFlatteningParameter 1 (and with the other of 2 and 3) end elsewhere,
    so a device or oracle that ignored the type would fail them."""
    z, g0, glob = load_case(path)
    own = glob["PPMFlatteningParameter"]
    assert own in (2, 3)
    last = int(z["nsteps"]) - 1
    for other in sorted({1, 2, 3} - {own}):
        g = g0.copy()
        gl = dict(glob, PPMFlatteningParameter=other)
        for s in range(int(z["nsteps"])):
            g.fill_periodic()
            g.dt = oracle.compute_time_step("port", g, gl, float(z["cfl"]))
            assert oracle.solve_ppm_de("port", g, gl, cycle=s)[0] == 1
        differ = sum(int(np.count_nonzero(g.fields[k] != z[f"out{last}_f{k}"]))
                     for k in range(g.nfields))
        assert differ > 100, (os.path.basename(path), other, differ)
