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


@pytest.mark.parametrize("graph", [True, False])
def test_level_comoving_matches_oracle(lib, graph):
    from enzo_ppm.level import Level
    from helpers import oracle_comoving_level
    grids = [problems.cosmological(n, seed=n)[0] for n in (24, 16, 20)]
    glob = problems.cosmological(4, seed=0)[1]
    model = comoving.EdS(0.1, 1.0)
    ref = [g.copy() for g in grids]
    _, dts = oracle_comoving_level(ref, glob, model, 1.0, 3, 0.4)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        lv.use_graph(graph)
        lv.set_comoving(True)
        t = 1.0
        for c in range(3):
            t, dt = comoving.level_step(lv, model, t, c, 0.4)
            assert dt == dts[c]
            assert lv.status()[0] == 1
        lv.download()
    finally:
        lv.close()
    for a, b in zip(lv.grids, ref):
        assert all(r[1] for r in field_report(a, b))


def test_level_single_grid_equals_per_grid(lib):
    """One cosmological grid as a level (graph replay) and as a device grid:
    the same comoving steps, bitwise."""
    from enzo_ppm.level import Level
    g, glob, _ = problems.cosmological(48, seed=9)
    model = comoving.EdS(0.1, 1.0)
    a = g.copy()
    dg = device.DeviceGrid(a, glob, lib=lib)
    lv = Level([g.copy()], glob, lib=lib)
    try:
        dg.set_comoving(True)
        lv.set_comoving(True)
        t1 = t2 = 1.0
        for c in range(3):
            t1, d1 = comoving.device_step(dg, model, t1, c, 0.4)
            t2, d2 = comoving.level_step(lv, model, t2, c, 0.4)
            assert d1 == d2
        assert dg.download() == 1
        lv.download()
    finally:
        dg.close()
        lv.close()
    assert all(r[1] for r in field_report(lv.grids[0], a))
