"""The reference's shock-tube answer tests on the GPU (run/Hydro/Hydro-1D:
Sod and Toro 1-7, 100 cells, outflow boundaries, gamma 1.4, CFL 0.9, run to
StopTime): the whole run device-resident -- outflow ghost fill
(hipPPMExternalBoundary), pressure and Courant step, rotated sweeps -- with
every step's dt and the end state bitwise equal to the oracle's run, and the
end state within the reference's own tolerances of the exact Riemann
solutions shipped with those tests (tests/golden)."""
import os

import numpy as np
import pytest

import oracle
from enzo_ppm import abi, device, shocktube
from helpers import field_report

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUTFLOW = [1] * 6


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


@pytest.mark.parametrize("name", sorted(shocktube.TESTS))
def test_shock_tube_device_resident(lib, name):
    dts = []

    def oracle_step(g, cycle):
        dts.append(g.dt)
        st, _ = oracle.solve_ppm_de("port", g, shocktube.GLOBALS, cycle=cycle)
        assert st == 1

    ref, ncyc = shocktube.run(name, oracle_step,
                              lambda g: oracle.compute_time_step("port", g, shocktube.GLOBALS,
                                                                 shocktube.CFL))
    g = shocktube.initial(name)
    t_stop = shocktube.TESTS[name][3]
    dg = device.DeviceGrid(g, shocktube.GLOBALS, lib=lib)
    try:
        t, cycle = 0.0, 0
        while t < t_stop * (1 - 1e-12):
            dg.external_boundary(OUTFLOW)
            dt = min(dg.pressure_and_dt(shocktube.CFL), t_stop - t)
            assert dt == dts[cycle], (cycle, dt, dts[cycle])
            dg.step(cycle, dt)
            t += dt
            cycle += 1
        dg.external_boundary(OUTFLOW)
        assert dg.download() == abi.SUCCESS
    finally:
        dg.close()
    assert cycle == ncyc
    rep = field_report(g, ref)
    assert all(r[1] for r in rep), rep
    for f, w, v in shocktube.compare(name, g, os.path.join(GOLD, shocktube.TESTS[name][4])):
        assert v <= 1.0, (name, f, w, v)
    assert np.all(g.d[g.active_slice()] > 0)
