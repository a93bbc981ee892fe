"""Parity of the HIP library (through the C ABI) with the oracle.

The oracle's C restatement is itself pinned bit-for-bit to the reference
Fortran (tests/test_oracle_reference.py and the golden fixtures); when the
reference build travels with the tree (oracle/_ref), it is compared directly
as well.  Bar: bitwise equality of every baryon field and captured face flux
(a maximum relative difference of 1e-11 after one step is also asserted,
and reported, so that a failure shows its size).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from enzo_ppm import abi, device, problems
from enzo_ppm.grid import SubgridFluxes
from helpers import bit_equal, evolved, field_report, fill_bc, max_rel, oracle_steps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


def _one_step(lib, g, glob, cfl, cycle, boxes=None, bc="periodic"):
    fill_bc(g, bc)
    g.dt = oracle.compute_time_step("port", g, glob, cfl)
    a, b = g.copy(), g.copy()
    fa = SubgridFluxes(a, boxes) if boxes else None
    fb = SubgridFluxes(b, boxes) if boxes else None
    st = device.solve_ppm_de(a, glob, cycle=cycle, fluxes=fa, lib=lib)
    assert st == abi.SUCCESS, (st, lib.hipPPMLastError())
    sto, fbk = oracle.solve_ppm_de("port", b, glob, cycle=cycle, fluxes=fb)
    assert sto == abi.SUCCESS
    return a, b, fa, fb, fbk


CASES = {
    "random_colours": lambda: problems.random_smooth([12, 10, 14], seed=1, ncolours=2),
    "gravity": lambda: problems.random_smooth([10, 12, 8], seed=2, ncolours=1, gravity=True),
    "no_dual_energy": lambda: problems.random_smooth([10, 12, 8], seed=3, dual_energy=False),
    "diffusion": lambda: problems.random_smooth([10, 12, 8], seed=4, ncolours=1, flags=(1, 1, 1)),
    "no_flatten_steepen": lambda: problems.random_smooth([16, 9, 11], seed=6, flags=(0, 0, 0)),
    "sod_2d": lambda: problems.sod_x(20, rank=2, gamma=1.4, dual_energy=True, flags=(1, 1, 1)),
    "turbulent_40": lambda: problems.turbulent(40, seed=3),
    "turbulent_mach30": lambda: problems.turbulent(24, seed=4, mach=30, sigma_lnrho=1.5),
    "stretched_grid": lambda: problems.random_smooth([12, 10, 11], seed=7, ncolours=1, stretch=0.2),
    "stretched_diffusion": lambda: problems.random_smooth([12, 10, 11], seed=7, ncolours=1,
                                                          stretch=0.2, flags=(1, 1, 1)),
    # the reference's longest pencil (fortran.def: MAX_ANY_SINGLE_DIRECTION
    # 1031 cells, ghosts included) along each axis, one active cell, and
    # ragged extents
    "longest_x": lambda: problems.random_smooth([1025, 3, 2], seed=11, ncolours=1),
    "longest_y": lambda: problems.random_smooth([3, 1025, 2], seed=11, ncolours=1),
    "longest_z": lambda: problems.random_smooth([2, 3, 1025], seed=11, ncolours=1),
    "one_cell": lambda: problems.random_smooth([1, 1, 1], seed=11, ncolours=1),
    "ragged": lambda: problems.random_smooth([5, 33, 2], seed=11, ncolours=1),
    "one_by_long": lambda: problems.random_smooth([1, 200, 1], seed=11, ncolours=1),
    # PPMFlatteningParameter 2 (Lagrangean, calcdiss.F:447-485) and 3
    # (multidimensional, :489-566) on strong shocks, where the three
    # flattenings differ (tests/test_golden.py checks the same on the
    # reference's fixtures): a Sod tube with diffusion, a Mach-30 box
    "flatten2_sod_diffusion": lambda: evolved(problems.sod_x(16, flags=(1, 2, 1)), 3),
    "flatten3_sod": lambda: evolved(problems.sod_x(16, flags=(1, 3, 0)), 3),
    "flatten3_mach30": lambda: evolved(_flatten(problems.turbulent(16, seed=4, mach=30,
                                                                   sigma_lnrho=1.5), 3), 2),
}


def _flatten(case, iflatten):
    g, glob, run = case
    return g, dict(glob, PPMFlatteningParameter=iflatten), run


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_one_step_matches_oracle(lib, name, cycle):
    g, glob, run = CASES[name]()
    a, b, _, _, _ = _one_step(lib, g, glob, run["cfl"], cycle)
    rep = field_report(a, b)
    worst = max(r[2] for r in rep)
    print(f"{name} cycle {cycle}: bitwise={all(r[1] for r in rep)} max_rel={worst:.3e}")
    assert worst <= 1e-11, rep
    assert all(r[1] for r in rep), rep


def test_config1_sod_64(lib):
    """64^3 Sod along x, diffusion/flattening/steepening on, one step (cycle 0)."""
    g, glob, run = problems.sod_x(64)
    a, b, _, _, _ = _one_step(lib, g, glob, run["cfl"], 0)
    rep = field_report(a, b)
    assert all(r[1] for r in rep), rep
    # mass conserved to round-off over the periodic box
    sl = a.active_slice()
    m0 = g.d[sl].sum()
    assert abs(a.d[sl].sum() - m0) <= 1e-12 * m0


@pytest.mark.parametrize("solver", [5, 1, 4], ids=["twoshock", "hll", "hllc"])
def test_subgrid_face_fluxes(lib, solver):
    g, glob, run = problems.random_smooth([20, 18, 16], seed=5, ncolours=2)
    glob = dict(glob, RiemannSolver=solver)
    boxes = [((7, 5, 4), (12, 9, 8)), ((3, 3, 3), (22, 20, 18)), ((10, 3, 3), (10, 20, 18))]
    for cycle in range(3):
        a, b, fa, fb, _ = _one_step(lib, g, glob, run["cfl"], cycle, boxes=boxes)
        assert all(r[1] for r in field_report(a, b))
        assert bit_equal(fa.flat(), fb.flat())
        assert np.count_nonzero(fa.flat()) > 0.9 * fa.flat().size


def test_rank1_sod(lib):
    g, glob, run = problems.sod_x(50, rank=1, gamma=1.4, dual_energy=False, flags=(0, 0, 0))
    for cycle in range(5):
        a, b, _, _, _ = _one_step(lib, g, glob, 0.9, cycle, bc="outflow")
        assert all(r[1] for r in field_report(a, b))
        g = a


def test_against_reference_build(lib):
    if not oracle.available("ref"):
        pytest.skip("oracle/_ref not built in this tree")
    g, glob, run = problems.turbulent(32, seed=8)
    g.fill_periodic()
    g.dt = oracle.compute_time_step("ref", g, glob, run["cfl"])
    a, b = g.copy(), g.copy()
    assert device.solve_ppm_de(a, glob, cycle=1, lib=lib) == abi.SUCCESS
    assert oracle.solve_ppm_de("ref", b, glob, cycle=1)[0] == abi.SUCCESS
    assert all(r[1] for r in field_report(a, b))


def test_device_resident_steps(lib):
    """Pressure + Courant pass, rotated sweeps and periodic ghost fill all on
    the device for several steps, against the oracle step loop."""
    g, glob, run = problems.turbulent(32, seed=11)
    ref = g.copy()
    dg = device.DeviceGrid(g, glob, lib=lib)
    nsteps = 6
    dts = []
    for n in range(nsteps):
        dg.periodic_boundary()
        dts.append(dg.pressure_and_dt(run["cfl"]))
        dg.step(n)
    assert dg.download() == abi.SUCCESS
    dg.close()
    for n in range(nsteps):
        fill_bc(ref, "periodic")
        ref.dt = oracle.compute_time_step("port", ref, glob, run["cfl"])
        assert ref.dt == dts[n], (n, ref.dt, dts[n])
        assert oracle.solve_ppm_de("port", ref, glob, cycle=n)[0] == abi.SUCCESS
    # the device state after the last step has no ghost fill; compare actives
    sl = g.active_slice()
    for x, y in zip(g.fields, ref.fields):
        assert bit_equal(x[sl], y[sl])


GOLD_STEPS = sorted(__import__("glob").glob(__import__("os").path.join(
    __import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden", "step_*.npz")))


@pytest.mark.parametrize("path", GOLD_STEPS, ids=[p.rsplit("/", 1)[-1] for p in GOLD_STEPS])
def test_device_matches_golden(lib, path):
    """The reference's own outputs (tests/golden, written by the flang build
    of its Fortran): the device path replays every step bitwise, face fluxes
    included; the HLL-solver fixtures run flux_hll over whole slices."""
    import test_golden as tg
    z, g, glob = tg.load_case(path)

    def step(g, s, fl):
        st = device.solve_ppm_de(g, glob, cycle=s, fluxes=fl, lib=lib)
        assert st == abi.SUCCESS, (st, lib.hipPPMLastError())

    tg.replay(z, g, glob, step, path.rsplit("/", 1)[-1])


HLL_CASES = ["turbulent_40", "diffusion", "gravity", "no_dual_energy", "sod_2d", "stretched_grid",
             "longest_hll", "pencil_200_diffusion", "pencil_400"]
CASES["longest_hll"] = lambda: problems.random_smooth([2, 1019, 3], seed=12, ncolours=1)
# pencils of 206 and 406 cells: the 256- and 512-wide one-pencil workgroups
CASES["pencil_200_diffusion"] = lambda: problems.random_smooth([2, 200, 3], seed=21, ncolours=1,
                                                               flags=(1, 1, 1))
CASES["pencil_400"] = lambda: problems.random_smooth([400, 3, 2], seed=22, ncolours=1)


@pytest.mark.parametrize("solver", [1, 4], ids=["hll", "hllc"])
@pytest.mark.parametrize("name", HLL_CASES)
@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_hll_solver_matches_oracle(lib, name, cycle, solver):
    """RiemannSolver HLL (typedefs.h: 1; Grid_xEulerSweep.C:245-256) and
    HLLC (4; :258-269): every interface takes flux_hll's / flux_hllc's
    fluxes; bitwise against the oracle."""
    g, glob, run = CASES[name]()
    glob = dict(glob, RiemannSolver=solver)
    a, b, _, _, fbk = _one_step(lib, g, glob, run["cfl"], cycle)
    assert fbk == 0
    rep = field_report(a, b)
    assert all(r[1] for r in rep), rep


def test_hll_solver_refuses_longer_pencils(lib):
    """The HLL solver's one-pencil workgroups hold 1025 cells: a longer
    pencil (up to the reference's 1031) is refused before anything runs."""
    g, glob, run = CASES["longest_x"]()
    glob = dict(glob, RiemannSolver=1)
    fill_bc(g, "periodic")
    g.dt = oracle.compute_time_step("port", g, glob, run["cfl"])
    a = g.copy()
    assert device.solve_ppm_de(a, glob, cycle=0, lib=lib) == abi.ERR_NOT_SUPPORTED
    assert b"HLL" in lib.hipPPMLastError()



@pytest.mark.parametrize("case", [("3d_z_ok", 2, 3, 3.0, 1, 4), ("3d_x_fails", 0, 0, 3.210948911011219, -3, 6)])
def test_hllc_fallback_matches_oracle(lib, case):
    """flux_hllc's per-cell flux_hll fallback (flux_hllc.F:344-391): status,
    fallback count and, when the step succeeds, every field as the oracle."""
    name, axis, seed, factor, status, nfb = case
    g, glob, _ = problems.two_stream(32, axis=axis, rank=3, seed=seed, ncolours=1)
    glob = dict(glob, RiemannSolver=4)
    g.dt = oracle.compute_time_step("port", g, glob, 0.4) * factor
    a, b = g.copy(), g.copy()
    assert oracle.solve_ppm_de("port", a, glob, cycle=axis) == (status, nfb)
    h = C.c_void_p(None)
    try:
        st = device.solve_ppm_de(b, glob, cycle=axis, handle=h, lib=lib)
        assert (st, lib.hipPPMFallbackCount(h)) == (status, nfb)
    finally:
        if h.value:
            lib.hipPPMDestroy(h)
    if status == 1:
        assert all(r[1] for r in field_report(b, a))


GOLD_FALLBACKS = sorted(__import__("glob").glob(__import__("os").path.join(
    __import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden", "fallback_*.npz")))


@pytest.mark.parametrize("path", GOLD_FALLBACKS, ids=[p.rsplit("/", 1)[-1] for p in GOLD_FALLBACKS])
def test_device_matches_golden_fallback(lib, path):
    """Per-cell fallback steps written by the reference (TwoShock and HLLC):
    the device's status, fallback count and fields, except the gas energy of
    the cells after fallback cells (read from unset locals there)."""
    from enzo_ppm.grid import Grid
    from helpers import fallback_differences
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
    h = C.c_void_p(None)
    try:
        st = device.solve_ppm_de(g, glob, cycle=int(z["cycle"]), handle=h, lib=lib)
        fb = lib.hipPPMFallbackCount(h)
    finally:
        if h.value:
            lib.hipPPMDestroy(h)
    assert (st, fb) == (int(z["status"]), int(z["fallbacks"]))
    fallback_differences(g, [z[f"out_f{k}"] for k in range(g.nfields)], fb)


@pytest.mark.parametrize("n", [24, 72, 170])
def test_host_round_trip(lib, n):
    """The host-array transfers: the direct path (24^3), the pinned staging
    ring with one chunk per field (72^3) and with two (170^3, 45 MB per
    field): every array back bit for bit, in both directions."""
    from helpers import check_host_round_trip
    check_host_round_trip(lib, n)


def test_periodic_fill(lib):
    """Device periodic ghost fill against the host fill, bitwise, per grid
    and on a level; narrow grids refused."""
    from helpers import check_periodic_fill
    check_periodic_fill(lib)
