"""Writes the golden fixtures of tests/golden from the reference itself.

The reference's hot-path Fortran (compiled by oracle/Makefile from
/root/reference into oracle/_ref) is run on small seeded inputs; each fixture
stores the inputs and the reference outputs as data (npz, float64 bit
patterns), so tests can pin the oracle and the GPU path where the reference
build is absent.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "enzo-dev_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from helpers import FALLBACK_CASES, fallback_case  # noqa: E402
from enzo_ppm import problems  # noqa: E402
from enzo_ppm.grid import SubgridFluxes  # noqa: E402

CASES = {
    # name: (factory, cfl, steps, boxes)
    "sod_diffusion_12": (lambda: problems.sod_x(12), 0.4, 2, None),
    "turbulent_10_c3": (lambda: problems.turbulent(10, seed=21), 0.4, 3, None),
    "gravity_8": (lambda: problems.random_smooth([8, 7, 9], seed=22, ncolours=1, gravity=True), 0.4, 2, None),
    "no_dual_8": (lambda: problems.random_smooth([9, 8, 7], seed=23, dual_energy=False), 0.4, 2, None),
    "diffusion_8": (lambda: problems.random_smooth([8, 9, 7], seed=24, ncolours=1, flags=(1, 1, 1)), 0.4, 2, None),
    "fluxes_10": (lambda: problems.random_smooth([10, 9, 8], seed=25, ncolours=1), 0.4, 1,
                  [((5, 4, 4), (8, 7, 6)), ((3, 3, 3), (12, 11, 10))]),
    "blast_10": (lambda: problems.amr_level(1, seed=26, nmin=8, nmax=12)[0][0], 0.4, 2, None),
    # RiemannSolver HLL (typedefs.h: 1): flux_hll over whole slices
    "hll_turbulent_10": (lambda: with_glob(problems.turbulent(10, seed=27), RiemannSolver=1),
                         0.4, 2, None),
    "hll_diffusion_8": (lambda: with_glob(problems.random_smooth([8, 9, 7], seed=28, ncolours=1,
                                                                flags=(1, 1, 1)), RiemannSolver=1),
                        0.4, 2, None),
    # PPMFlatteningParameter 2 (calcdiss.F:447-485) and 3 (:489-566) on
    # strong shocks, where the three flattenings differ (the smooth fields
    # above leave every flattening at zero)
    "sod_flatten2_12": (lambda: with_glob(problems.sod_x(12), PPMFlatteningParameter=2),
                        0.4, 4, None),
    "sod_flatten3_12": (lambda: with_glob(problems.sod_x(12), PPMFlatteningParameter=3),
                        0.4, 4, None),
    "mach30_flatten3_10": (lambda: with_glob(problems.turbulent(10, seed=31, mach=30,
                                                                sigma_lnrho=1.5),
                                             PPMFlatteningParameter=3), 0.4, 3, None),
    # RiemannSolver HLLC (4): flux_hllc over whole slices
    "hllc_turbulent_10": (lambda: with_glob(problems.turbulent(10, seed=29), RiemannSolver=4),
                          0.4, 2, None),
    "hllc_gravity_8": (lambda: with_glob(problems.random_smooth([8, 7, 9], seed=30, ncolours=1,
                                                               gravity=True), RiemannSolver=4),
                       0.4, 2, None),
}


def with_glob(case, **kw):
    g, glob, run = case
    return g, dict(glob, **kw), run


def grid_arrays(g):
    return {f"f{k}": f for k, f in enumerate(g.fields)}


def make(name):
    fac, cfl, steps, boxes = CASES[name]
    r = fac()
    if isinstance(r, tuple):
        g, glob, _ = r
    else:
        g = r
        glob = dict(Gamma=5 / 3, PPMSteepeningParameter=1, PPMFlatteningParameter=1,
                    PPMDiffusionParameter=0, DualEnergyFormalism=1)
    out = {"glob_keys": np.array(sorted(glob)), "glob_vals": np.array([float(glob[k]) for k in sorted(glob)]),
           "active": np.array(g.active), "rank": g.rank, "ncolours": g.ncolours,
           "dual": int(g.dual_energy), "gravity": int(g.gravity), "dx": np.array(g.dx),
           "nsteps": steps, "cfl": cfl}
    if g.gravity:
        for a in range(3):
            out[f"accel{a}"] = g.accel[a]
    periodic = not name.startswith("blast")
    for s in range(steps):
        if periodic:
            g.fill_periodic()
        g.dt = oracle.compute_time_step("ref", g, glob, cfl)
        out[f"dt{s}"] = g.dt
        if s == 0:
            for k, f in enumerate(g.fields):
                out[f"in0_f{k}"] = f.copy()
        fl = SubgridFluxes(g, boxes) if boxes else None
        st, fb = oracle.solve_ppm_de("ref", g, glob, cycle=s, fluxes=fl)
        assert st == 1 and fb == 0, (name, st, fb)
        for k, f in enumerate(g.fields):
            out[f"out{s}_f{k}"] = f.copy()
        if fl is not None:
            out[f"flux{s}"] = fl.flat()
            out["boxes"] = np.array(boxes)
    np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **out)


def make_twoshock():
    import ctypes as C
    rng = np.random.default_rng(31)
    idim, jdim = 512, 1
    dls, drs = rng.uniform(0.05, 4, (2, jdim, idim))
    pls, prs = np.exp(rng.uniform(-5, 3, (2, jdim, idim)))
    uls, urs = rng.uniform(-3, 3, (2, jdim, idim))
    pbar, ubar = np.zeros((jdim, idim)), np.zeros((jdim, idim))
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    i = lambda v: C.byref(C.c_int64(v))
    d = lambda v: C.byref(C.c_double(v))
    pl, pr = pls.copy(), prs.copy()
    oracle.lib("ref").twoshock_(p(dls), p(drs), p(pl), p(pr), p(uls), p(urs), i(idim), i(jdim),
                                i(1), i(idim), i(1), i(1), d(0.0), d(5 / 3), d(1e-20), i(0),
                                p(pbar), p(ubar), i(0), p(pbar), i(1), d(1e-3))
    np.savez_compressed(os.path.join(HERE, "twoshock_512.npz"), dls=dls, drs=drs, pls=pls, prs=prs,
                        uls=uls, urs=urs, pbar=pbar, ubar=ubar)


def make_hllc_fallback():
    """One SolvePPM_DE call with RiemannSolver HLLC in which four cells take
    flux_hllc's per-cell flux_hll fallback (flux_hllc.F:344-391): a two-stream
    3-D grid streaming along z, three Courant steps long."""
    g, glob, _ = problems.two_stream(32, axis=2, rank=3, seed=3, ncolours=1)
    glob = dict(glob, RiemannSolver=4)
    g.dt = oracle.compute_time_step("ref", g, glob, 0.4) * 3.0
    out = {"glob_keys": np.array(sorted(glob)), "glob_vals": np.array([float(glob[k]) for k in sorted(glob)]),
           "active": np.array(g.active), "rank": g.rank, "ncolours": g.ncolours, "dx": np.array(g.dx),
           "dt": g.dt, "cycle": 2}
    for d in range(3):
        out[f"cellwidth{d}"] = g.cellwidth[d].copy()
    for k, f in enumerate(g.fields):
        out[f"in_f{k}"] = f.copy()
    st, fb = oracle.solve_ppm_de("ref", g, glob, cycle=2)
    assert st == 1 and fb > 0, (st, fb)
    out["status"], out["fallbacks"] = st, fb
    for k, f in enumerate(g.fields):
        out[f"out_f{k}"] = f.copy()
    np.savez_compressed(os.path.join(HERE, "fallback_hllc_3d_z_ok.npz"), **out)


def make_fallback(name):
    """One SolvePPM_DE call of a per-cell fallback case (tests/helpers.py)
    with the reference's status, fallback count and fields.  The gas energy
    of the cell after each fallback cell is the one value the reference
    computes from locals it never set (flux_hll.F:304-307 with i1 = i2)."""
    g, glob, cycle = fallback_case(name)
    out = {"glob_keys": np.array(sorted(glob)), "glob_vals": np.array([float(glob[k]) for k in sorted(glob)]),
           "active": np.array(g.active), "rank": g.rank, "ncolours": g.ncolours, "dx": np.array(g.dx),
           "dt": g.dt, "cycle": cycle}
    for d in range(3):
        out[f"cellwidth{d}"] = g.cellwidth[d].copy()
    for k, f in enumerate(g.fields):
        out[f"in_f{k}"] = f.copy()
    st, fb = oracle.solve_ppm_de("ref", g, glob, cycle=cycle)
    assert st == FALLBACK_CASES[name][-1] and fb > 0, (name, st, fb)
    out["status"], out["fallbacks"] = st, fb
    for k, f in enumerate(g.fields):
        out[f"out_f{k}"] = f.copy()
    np.savez_compressed(os.path.join(HERE, f"fallback_{name}.npz"), **out)


if __name__ == "__main__":
    if not oracle.available("ref"):
        oracle.build("ref")
    only = sys.argv[1:]          # case names to (re)write; all by default
    for n in CASES:
        if only and n not in only:
            continue
        make(n)
        print("wrote", n)
    if not only or "hllc_fallback" in only:
        make_hllc_fallback()
        print("wrote hllc fallback")
    if only:
        sys.exit(0)
    make_twoshock()
    print("wrote twoshock")
    for n in ("1d_ok", "3d_y_ok", "3d_z_stretched_ok"):
        make_fallback(n)
        print("wrote fallback", n)
