"""Comoving steps on the GPU (hipPPMSetComoving, hipPPMSetExpansionFactors,
hipPPMComovingExpansionTerms) against the oracle, bitwise: cosmological
boxes with lognormal density (sigma_lnrho 2), Mach-100 velocities, a static
acceleration field and six species, in an Einstein-de Sitter expansion from
a = 0.1."""
import pytest

from enzo_ppm import abi, comoving, device, problems
from helpers import field_report, oracle_comoving_steps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


@pytest.mark.parametrize("n,dual,steps", [(32, 1, 4), (24, 0, 3), (40, 1, 2)])
def test_comoving_steps_match_oracle(lib, n, dual, steps):
    g, glob, _ = problems.cosmological(n, seed=n + dual)
    glob = dict(glob, DualEnergyFormalism=dual)
    model = comoving.EdS(0.1, 1.0)
    ref = g.copy()
    _, dts = oracle_comoving_steps("port", ref, glob, model, 1.0, steps, 0.4)
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        dg.set_comoving(True)
        t = 1.0
        for c in range(steps):
            t, dt = comoving.device_step(dg, model, t, c, 0.4)
            assert dt == dts[c], (c, dt, dts[c])
        assert dg.download() == 1
    finally:
        dg.close()
    rep = field_report(g, ref)
    assert all(r[1] for r in rep), rep


def test_comoving_off_again_is_plain(lib):
    """Switching comoving handling off restores the plain step bitwise."""
    import oracle
    g, glob, _ = problems.cosmological(16, seed=3)
    ref = g.copy()
    ref.fill_periodic()
    ref.dt = oracle.compute_time_step("port", ref, glob, 0.4)
    assert oracle.solve_ppm_de("port", ref, glob, cycle=0)[0] == 1
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        dg.set_comoving(True)
        dg.set_expansion_factors(0.1, 0.1001)
        dg.set_comoving(False)
        dg.periodic_boundary()
        dg.pressure_and_dt(0.4)
        dg.step(0)
        assert dg.download() == 1
    finally:
        dg.close()
    assert all(r[1] for r in field_report(g, ref))
