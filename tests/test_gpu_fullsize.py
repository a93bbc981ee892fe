"""Full-size parity: the 256^3 turbulent workload (lognormal density
sigma_lnrho 0.5, k^-2 Mach-3 velocities, 3 colours, dual energy, periodic)
for 20 device-resident steps, compared with the oracle step by step:
- the plain start: every step's Courant dt and every field bitwise;
- the learned start evaluated at every interface (gate 0): after one step
  max relative difference <= 1e-11 in density and energy (per cell) and in
  momentum (max |d(rho v)| over max |rho v|: a per-cell ratio is undefined
  where v crosses 0), after 20 steps L1 relative difference <= 1e-9 per
  field;
and mass conserved over the periodic box to round-off."""
import numpy as np
import pytest

import learn_checks as lc
import oracle
from enzo_ppm import abi, device, learn, problems
from helpers import bit_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


def l1(a, b):
    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def test_turbulent_256_twenty_steps_step_by_step(lib):
    # the network, trained on 64^3 states of both generators
    g64, glob64, _ = problems.sod_x(64)
    r1, _ = learn.harvest(g64, glob64, steps=1, lib=lib)
    g64, glob64, _ = problems.turbulent(64, seed=5)
    r2, _ = learn.harvest(g64, glob64, steps=1, lib=lib)
    w, _ = learn.train(np.concatenate([r1, r2]), lc.GAMMA, epochs=4, batch=4096, lib=lib)
    w[learn.OFF_GATE] = 0.0
    n, steps = 256, 20
    g, glob, run = problems.turbulent(n, seed=2024)
    ref, gl = g.copy(), g.copy()
    sl = g.active_slice()
    mass0 = ref.fields[0][sl].sum()
    dg = device.DeviceGrid(g, glob, lib=lib)
    dl = device.DeviceGrid(gl, glob, lib=lib)
    try:
        dl.set_learned(w)
        for s in range(steps):
            for d in (dg, dl):
                d.periodic_boundary()
            dt = dg.pressure_and_dt(run["cfl"])
            dl.pressure_and_dt(run["cfl"])
            dg.step(s)
            dl.step(s)
            assert dg.download() == abi.SUCCESS and dl.download() == abi.SUCCESS
            ref.fill_periodic()
            ref.dt = oracle.compute_time_step("port", ref, glob, run["cfl"])
            assert dt == ref.dt, (s, dt, ref.dt)
            st, _ = oracle.solve_ppm_de("port", ref, glob, cycle=s, nthreads=16)
            assert st == abi.SUCCESS
            for k, (x, y) in enumerate(zip(g.fields, ref.fields)):
                assert bit_equal(x[sl], y[sl]), (s, k)
            if s == 0:
                d0 = ref.fields[0][sl]
                for k in (0, 1):        # density, total energy
                    rel = np.max(np.abs(gl.fields[k][sl] - ref.fields[k][sl]) /
                                 np.abs(ref.fields[k][sl]))
                    assert rel <= 1e-11, (k, rel)
                for k in (2, 3, 4):     # momentum, relative to |rho v| scale
                    mom = gl.fields[0][sl] * gl.fields[k][sl]
                    mref = d0 * ref.fields[k][sl]
                    rel = np.max(np.abs(mom - mref)) / np.max(np.abs(mref))
                    assert rel <= 1e-11, (k, rel)
            err = [l1(x[sl], y[sl]) for x, y in zip(gl.fields, ref.fields)]
            print(f"step {s}: dt {dt:.6e} plain bitwise; learned L1 max {max(err):.2e}", flush=True)
        assert max(err) <= 1e-9, err
    finally:
        dg.close()
        dl.close()
    for h in (g, gl):
        mass = h.fields[0][sl].sum()
        assert abs(mass - mass0) <= 1e-12 * mass0, (mass, mass0)


def test_amr_level_512_grids_matches_oracle(lib):
    """The AMR-level workload at full size: 512 grids, axes log-uniform in
    [8, 64] (9.44 M active cells), Sedov blasts in sinusoidal flow, batched
    (graph replay, default tile shapes) for two steps at the level's Courant
    dt, against the oracle's per-grid loop: every dt and every field of every
    grid bitwise."""
    from enzo_ppm.level import Level
    from helpers import level_reference
    grids, glob, _ = problems.amr_level(512, seed=7, nmin=8, nmax=64, ncolours=0)
    ref = [g.copy() for g in grids]
    dts = level_reference(ref, glob, 2)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        for s in range(2):
            lv.step(s, 0.0, 0.4)
            st, dt, _ = lv.status()
            assert st == 1 and dt == dts[s], (s, st, dt, dts[s])
        lv.download()
        for k, (a, b) in enumerate(zip(lv.grids, ref)):
            for x, y in zip(a.fields, b.fields):
                assert bit_equal(x, y), k
    finally:
        lv.close()
