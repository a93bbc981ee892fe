"""The library's own HIP source, built for the CPU (tests/emu/hipcpu), against
the oracle through the same C ABI the GPU tests use.

This runs the library's host orchestration -- allocations, launch shapes,
sweep order, slice-serial diffusion launches, fallback lists -- and its
kernels work-item by work-item, with every device buffer against a guard
page, so an out-of-bounds index stops the test.  Bar: bitwise equality with
the oracle, and its status and fallback count.
"""
import ctypes as C

import pytest

import oracle
from enzo_ppm import device, problems
from enzo_ppm.grid import SubgridFluxes
from helpers import (FALLBACK_CASES, bit_equal, cpu_library, evolved, fallback_case, field_report,
                     fill_bc)


@pytest.fixture(scope="module")
def cpulib():
    return cpu_library()


def run(lib, g, glob, cycle, fluxes=None):
    h = C.c_void_p(None)
    try:
        st = device.solve_ppm_de(g, glob, cycle=cycle, fluxes=fluxes, handle=h, lib=lib)
        return st, lib.hipPPMFallbackCount(h)
    finally:
        if h.value:
            lib.hipPPMDestroy(h)


# guard after every buffer for all cases; before every buffer for a few
GUARDED = [(n, "high") for n in sorted(FALLBACK_CASES)] + [
    (n, "low") for n in ("1d_ok", "3d_y_diffusion_hll_fails", "3d_z_stretched_ok")]


@pytest.mark.parametrize("name,guard", GUARDED)
def test_fallback_cases(cpulib, name, guard, monkeypatch):
    monkeypatch.setenv("HIPCPU_GUARD", guard)
    g, glob, cycle = fallback_case(name)
    a, b = g.copy(), g.copy()
    so, fo = oracle.solve_ppm_de("port", a, glob, cycle=cycle)
    assert (so, fo > 0) == (FALLBACK_CASES[name][-1], True)
    assert run(cpulib, b, glob, cycle) == (so, fo)
    if so == 1:
        assert all(r[1] for r in field_report(b, a))


CASES = {
    "diffusion": lambda: problems.random_smooth([10, 12, 8], seed=4, ncolours=1, flags=(1, 1, 1)),
    "stretched_diffusion": lambda: problems.random_smooth([12, 10, 11], seed=7, ncolours=1,
                                                          stretch=0.2, flags=(1, 1, 1)),
    "gravity": lambda: problems.random_smooth([10, 12, 8], seed=2, ncolours=1, gravity=True),
    "sod_2d": lambda: problems.sod_x(20, rank=2, gamma=1.4, flags=(1, 1, 1)),
    "turbulent": lambda: problems.turbulent(12, seed=3),
    # the longest pencil the reference allows (MAX_ANY_SINGLE_DIRECTION 1031
    # cells, ghosts included), one active cell, ragged extents
    "longest_y": lambda: problems.random_smooth([3, 1025, 2], seed=11, ncolours=1),
    "one_cell": lambda: problems.random_smooth([1, 1, 1], seed=11, ncolours=1),
    "ragged": lambda: problems.random_smooth([5, 33, 2], seed=11, ncolours=1),
    # flattening types 2 and 3 (calcdiss.F:447-566) on strong shocks, where
    # the three flattenings differ
    "flatten2_sod_diffusion": lambda: evolved(problems.sod_x(12, flags=(1, 2, 1)), 3),
    "flatten3_sod": lambda: evolved(problems.sod_x(12, flags=(1, 3, 0)), 3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_steps(cpulib, name):
    g, glob, _ = CASES[name]()
    for cycle in (0, 2):
        fill_bc(g, "periodic")
        g.dt = oracle.compute_time_step("port", g, glob, 0.4)
        a, b = g.copy(), g.copy()
        assert oracle.solve_ppm_de("port", a, glob, cycle=cycle)[0] == 1
        assert run(cpulib, b, glob, cycle) == (1, 0)
        assert all(r[1] for r in field_report(b, a))
        g = a


@pytest.mark.parametrize("solver", [5, 1, 4], ids=["twoshock", "hll", "hllc"])
def test_face_fluxes(cpulib, solver):
    g, glob, _ = problems.random_smooth([20, 18, 16], seed=5, ncolours=2)
    glob = dict(glob, RiemannSolver=solver)
    boxes = [((7, 5, 4), (12, 9, 8)), ((3, 3, 3), (22, 20, 18)), ((10, 3, 3), (10, 20, 18))]
    g.fill_periodic()
    g.dt = oracle.compute_time_step("port", g, glob, 0.4)
    a, b = g.copy(), g.copy()
    fa, fb = SubgridFluxes(a, boxes), SubgridFluxes(b, boxes)
    assert oracle.solve_ppm_de("port", a, glob, cycle=1, fluxes=fa)[0] == 1
    assert run(cpulib, b, glob, 1, fluxes=fb)[0] == 1
    assert bit_equal(fa.flat(), fb.flat())
    assert all(r[1] for r in field_report(b, a))


def test_device_resident_steps(cpulib):
    """Boundary fill, pressure + Courant pass and rotated sweeps on the
    device state, three steps, against the oracle with the same time steps."""
    g, glob, _ = problems.turbulent(12, seed=8)
    a = g.copy()
    dg = device.DeviceGrid(g, glob, lib=cpulib)
    try:
        for cycle in range(3):
            dg.periodic_boundary()
            dt = dg.pressure_and_dt(0.4)
            dg.step(cycle)
            a.fill_periodic()
            a.dt = oracle.compute_time_step("port", a, glob, 0.4)
            assert dt == a.dt
            assert oracle.solve_ppm_de("port", a, glob, cycle=cycle)[0] == 1
        assert dg.download() == 1
    finally:
        dg.close()
    assert all(r[1] for r in field_report(g, a))


@pytest.mark.parametrize("dual", [1, 0])
def test_comoving_steps(cpulib, dual):
    """Comoving steps on the device state (Courant with a(t), sweeps on
    a(t+dt/2)-scaled widths, expansion terms with the time-centred pressure)
    against the oracle, bitwise, with gravity and 6 species."""
    from enzo_ppm import comoving
    from helpers import oracle_comoving_steps
    g, glob, _ = problems.cosmological(12, seed=4)
    glob = dict(glob, DualEnergyFormalism=dual)
    model = comoving.EdS(0.1, 1.0)
    ref = g.copy()
    _, dts = oracle_comoving_steps("port", ref, glob, model, 1.0, 3, 0.4)
    dg = device.DeviceGrid(g, glob, lib=cpulib)
    try:
        dg.set_comoving(True)
        t = 1.0
        for c in range(3):
            t, dt = comoving.device_step(dg, model, t, c, 0.4)
            assert dt == dts[c]
        assert dg.download() == 1
    finally:
        dg.close()
    assert all(r[1] for r in field_report(g, ref))


# ---- batched AMR level (hipPPMLevel*) --------------------------------------
from enzo_ppm.level import Level  # noqa: E402
from helpers import check_tune_times, level_grids, level_reference, sod_level  # noqa: E402


def level_check(lib, grids, glob, steps, shapes=None, graph=True, dt=None, persistent=False):
    ref = [g.copy() for g in grids]
    dts = level_reference(ref, glob, steps, dt=dt)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        lv.use_graph(graph)
        assert lv.persistent(persistent) == persistent
        if shapes is not None:
            for i, g in enumerate(lv.grids):
                for d in range(g.rank):
                    lv.set_shape(i, d, shapes)
            lv.commit_shapes()
        for s in range(steps):
            lv.step(s, dt if dt is not None else 0.0, 0.4)
            st, dtg, nfb = lv.status()
            assert st == 1 and dtg == dts[s], (st, dtg, dts[s])
        lv.download()
        for a, b in zip(lv.grids, ref):
            rep = field_report(a, b)
            assert all(r[1] for r in rep), (a.active, rep)
    finally:
        lv.close()


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_level_blast_shapes(cpulib, shape):
    grids, glob = level_grids("blast", 4)
    level_check(cpulib, grids, glob, 3, shapes=shape)


def test_level_direct_launches(cpulib):
    grids, glob = level_grids("blast", 3, seed=9)
    level_check(cpulib, grids, glob, 2, graph=False)


@pytest.mark.parametrize("kind", ["gravity", "stretched", "diffusion"])
def test_level_kinds(cpulib, kind):
    grids, glob = level_grids(kind, 3)
    level_check(cpulib, grids, glob, 2)


@pytest.mark.parametrize("iflatten", [2, 3])
def test_level_flattening_types(cpulib, iflatten):
    """PPMFlatteningParameter 2 and 3 (calcdiss.F:447-566) on the batched
    level: a level of Sod tubes (strong shocks in every slice's first line,
    where each slice's flattening is taken; diffusion on), every grid
    bitwise as the oracle's, whose result differs from flattening type 1's."""
    grids, glob = sod_level(iflatten)
    level_check(cpulib, grids, glob, 3)
    a, b = [g.copy() for g in grids], [g.copy() for g in grids]
    level_reference(a, glob, 3)
    level_reference(b, dict(glob, PPMFlatteningParameter=1), 3)
    assert any(not bit_equal(x, y) for ga, gb in zip(a, b) for x, y in zip(ga.fields, gb.fields))


@pytest.mark.parametrize("kind", ["blast", "diffusion"])
@pytest.mark.parametrize("graph", [True, False])
def test_level_persistent(cpulib, kind, graph):
    """Persistent sweep launches (workgroups pulling work items from per-XCD
    counters, hipPPMLevelPersistent): every grid bitwise as the oracle's
    per-grid loop, over steps that reuse the counters."""
    grids, glob = level_grids(kind, 4)
    level_check(cpulib, grids, glob, 3, persistent=True, graph=graph)


def test_level_fallback(cpulib):
    """A level whose grids need the per-cell fallback: same fields and the
    summed fallback count of the per-grid oracle calls."""
    grids = []
    for name in ("3d_x_ok", "3d_x_stretched_ok"):
        g, glob, cycle = fallback_case(name)
        grids.append(g)
    dt = min(g.dt for g in grids)
    # both at the smaller dt: recompute the per-grid fallback counts there
    ref = [g.copy() for g in grids]
    nfb = 0
    for g in ref:
        g.dt = dt
        st, n = oracle.solve_ppm_de("port", g, glob, cycle=0)
        assert st == 1
        nfb += n
    lv = Level([g.copy() for g in grids], glob, lib=cpulib)
    try:
        lv.step(0, dt, 0.4)
        st, dtg, n = lv.status()
        assert (st, n) == (1, nfb)
        lv.download()
        for a, b in zip(lv.grids, ref):
            assert all(r[1] for r in field_report(a, b))
    finally:
        lv.close()


def test_level_tune_restores_state(cpulib):
    grids, glob = level_grids("blast", 3, seed=11)
    dt = min(oracle.compute_time_step("port", g, glob, 0.4) for g in grids)
    lv = Level([g.copy() for g in grids], glob, lib=cpulib)
    try:
        import numpy as np
        assert np.all(np.isnan(lv.tune_times()[0]))  # nothing timed yet
        choice = lv.tune(0, dt, repeats=1)
        assert all(0 <= c <= 2 for c in choice)
        check_tune_times(lv, choice)
        lv.step(0, dt, 0.4)
        assert lv.status()[0] == 1
        lv.download()
    finally:
        lv.close()
    ref = [g.copy() for g in grids]
    level_reference(ref, glob, 1, dt=dt)
    for a, b in zip(lv.grids, ref):
        assert all(r[1] for r in field_report(a, b))


@pytest.mark.parametrize("graph", [True, False])
def test_level_comoving(cpulib, graph):
    """A comoving level (two cosmological grids, periodic fill on the device,
    level Courant step with a(t), a(t+dt/2)-scaled widths, expansion terms)
    against the oracle's per-grid loop, bitwise."""
    from enzo_ppm import comoving
    from helpers import oracle_comoving_level
    grids = [problems.cosmological(n, seed=n)[0] for n in (12, 10)]
    glob = problems.cosmological(4, seed=0)[1]
    model = comoving.EdS(0.1, 1.0)
    ref = [g.copy() for g in grids]
    _, dts = oracle_comoving_level(ref, glob, model, 1.0, 3, 0.4)
    lv = Level([g.copy() for g in grids], glob, lib=cpulib)
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


def _level_download(lv):
    lv.download()
    return lv.grids


@pytest.mark.parametrize("parts,rank", [((2, 2, 2), 3), ((3, 2, 1), 2)])
def test_level_sibling_copies_match_oracle(cpulib, parts, rank):
    """hipPPMLevelCopyZones against the oracle's CopyZonesFromGrid restatement,
    every field, ghosts included, bitwise."""
    import numpy as np
    import boundary_oracle as bo
    n = [12, 12, 12][:rank]
    g, glob, _ = problems.random_smooth(n, seed=3, ncolours=1, rank=rank)
    g.fill_periodic()
    pieces, starts = bo.split_domain(g, parts)
    rng = np.random.default_rng(1)
    for h in pieces:              # ghosts hold junk before the copies
        sl = h.active_slice()
        for f in h.fields:
            act = f[sl].copy()
            f[...] = rng.normal(size=f.shape)
            f[sl] = act
    ref = [h.copy() for h in pieces]
    bo.copy_zones(ref, starts, g.active, periodic=True)
    lv = Level([h.copy() for h in pieces], glob, lib=cpulib)
    try:
        lv.set_siblings(starts, g.active, periodic=True)
        lv.copy_zones()
        lv.download()
    finally:
        lv.close()
    # the device holds the fields the hydro step uses: no Vel3 in 2-D
    held = [k for k in range(g.nfields) if not (rank < 3 and k == g.Vel3Num)]
    for a, b in zip(lv.grids, ref):
        rep = field_report(a, b)
        assert all(rep[k][1] for k in held), rep


def test_level_decomposed_box_equals_whole(cpulib):
    """A periodic 24^3 box as 8 sibling grids with device ghost copies before
    every level step equals the undivided box, bitwise in the active cells,
    with the same level dt.  Flattening and dual energy are off: in the
    reference both make a grid's result depend on where its edges are
    (flattening uses the coefficients of each slice's first line;
    pgas2d_dual's neighbour test reads the stale total energy of the
    sweep-direction ghost cells)."""
    import boundary_oracle as bo
    g, glob, _ = problems.turbulent(24, seed=6, ncolours=1, flags=(1, 0, 0))
    glob = dict(glob, DualEnergyFormalism=0)
    g.fill_periodic()
    pieces, starts = bo.split_domain(g, (2, 2, 2))
    whole = g.copy()
    lv = Level(pieces, glob, lib=cpulib)
    try:
        lv.set_siblings(starts, g.active, periodic=True)
        for s in range(2):
            lv.copy_zones()
            lv.step(s, 0.0, 0.4)
            st, dt, _ = lv.status()
            assert st == 1
            whole.fill_periodic()
            whole.dt = oracle.compute_time_step("port", whole, glob, 0.4)
            assert whole.dt == dt
            assert oracle.solve_ppm_de("port", whole, glob, cycle=s)[0] == 1
        lv.download()
    finally:
        lv.close()
    gh = g.ghost
    for h, st0 in zip(lv.grids, starts):
        lo = [st0[d] + gh for d in range(3)]
        w = tuple(slice(lo[d] + gh, lo[d] + gh + h.active[d]) for d in (2, 1, 0))
        sl = h.active_slice()
        for x, y in zip(h.fields, whole.fields):
            assert bit_equal(x[sl], y[w])


@pytest.mark.parametrize("name", ["interior", "undo", "on_left_x_edge", "on_right_z_edge"])
def test_flux_correction_matches_oracle(cpulib, name):
    import fluxcorr_checks as fc
    assert fc.run_case(cpulib, name) > 0


def test_level_flux_correction_matches_oracle(cpulib):
    """hipPPMLevelCorrectForRefinedFluxes on grid 1 of a two-grid level."""
    import fluxcorr_checks as fc
    fc.run_level_case(cpulib, "undo")


@pytest.mark.parametrize("r", [2, 4])
def test_flux_projection_matches_oracle(cpulib, r):
    import fluxcorr_checks as fc
    fc.run_projection(cpulib, r)


@pytest.mark.parametrize("name", ["left", "right", "left_undo_density", "right_undo_energy",
                                  "left_periodic", "right_periodic"])
def test_subling_flux_correction_matches_oracle(cpulib, name):
    import fluxcorr_checks as fc
    assert fc.run_subling_case(cpulib, name) > 0


@pytest.mark.parametrize("name", ["reflecting", "outflow", "mixed_inflow", "shearing_periodic",
                                  "inside_domain", "one_active_cell", "rank2"])
def test_external_boundary_matches_oracle(cpulib, name):
    import boundary_checks as bc
    assert bc.run_external_case(cpulib, name) > 0


def test_external_boundary_refusals_write_nothing(cpulib):
    """Face types outside the restated ones (5, the hydrostatic type) and a
    shearing face on a grid with fewer active cells than ghost layers are
    refused before any face is filled."""
    import ctypes as C
    from enzo_ppm import abi
    g, glob, _ = problems.random_smooth([2, 5, 4], seed=1, ncolours=1)
    a = g.copy()
    dg = device.DeviceGrid(a, glob, lib=cpulib)
    try:
        for types in ([1, 1, 5, 1, 1, 1], [4, 4, 1, 1, 1, 1]):
            st = cpulib.hipPPMExternalBoundary(dg.handle, C.byref(dg.para), abi.iarr(types, 6),
                                               abi.iarr([0, 0, 0], 3), abi.iarr(g.dims, 3),
                                               g.nfields, None)
            assert st == abi.ERR_NOT_SUPPORTED
        st = cpulib.hipPPMExternalBoundary(dg.handle, C.byref(dg.para), abi.iarr([2] * 6, 6),
                                           abi.iarr([0, 0, 0], 3), abi.iarr(g.dims, 3),
                                           g.nfields, None)
        assert st == abi.ERR_ARGUMENT          # inflow faces need their values
        assert dg.download() == 1
    finally:
        dg.close()
    for x, y in zip(a.fields, g.fields):
        assert bit_equal(x, y)


@pytest.mark.parametrize("graph", [True, False])
def test_level_subgrid_fluxes(cpulib, graph):
    """Face fluxes of subgrids recorded by the batched level equal the
    oracle's per-grid SolvePPM_DE with the same SubgridFluxes."""
    from helpers import level_flux_check
    level_flux_check(cpulib, "blast", graph=graph)


@pytest.mark.parametrize("types", [[0, 1, 3, 3, 2, 0], [1, 1, 0, 0, 4, 4]])
def test_level_external_boundary_matches_oracle(cpulib, types):
    import boundary_checks as bc
    bc.run_level_external(cpulib, types)


HLL_CASES = {
    "turbulent": lambda: problems.turbulent(12, seed=5),
    "diffusion": lambda: problems.random_smooth([10, 12, 8], seed=4, ncolours=1, flags=(1, 1, 1)),
    "gravity": lambda: problems.random_smooth([10, 12, 8], seed=2, ncolours=1, gravity=True),
    "no_dual_energy": lambda: problems.random_smooth([10, 12, 8], seed=3, dual_energy=False),
    "sod_2d": lambda: problems.sod_x(20, rank=2, gamma=1.4, flags=(1, 1, 1)),
    # pencils of 206 and 406 cells: the 256- and 512-wide one-pencil workgroups
    "pencil_200_diffusion": lambda: problems.random_smooth([2, 200, 3], seed=21, ncolours=1,
                                                           flags=(1, 1, 1)),
    "pencil_400": lambda: problems.random_smooth([400, 3, 2], seed=22, ncolours=1),
}


@pytest.mark.parametrize("solver", [1, 4], ids=["hll", "hllc"])
@pytest.mark.parametrize("name", sorted(HLL_CASES))
def test_hll_solver_steps(cpulib, name, solver):
    """RiemannSolver HLL (1: flux_hll over whole slices) and HLLC (4:
    flux_hllc): the device path equals the oracle bitwise over steps of
    every sweep order, with no fallback."""
    g, glob, _ = HLL_CASES[name]()
    glob = dict(glob, RiemannSolver=solver)
    for cycle in (0, 1, 2):
        fill_bc(g, "periodic")
        g.dt = oracle.compute_time_step("port", g, glob, 0.4)
        a, b = g.copy(), g.copy()
        assert oracle.solve_ppm_de("port", a, glob, cycle=cycle) == (1, 0)
        assert run(cpulib, b, glob, cycle) == (1, 0)
        assert all(r[1] for r in field_report(b, a))
        g = a



@pytest.mark.parametrize("case", [("3d_z_ok", 2, 3, 3.0, 1, 4), ("3d_x_fails", 0, 0, 3.210948911011219, -3, 6)])
def test_hllc_fallback(cpulib, case):
    """flux_hllc's per-cell flux_hll fallback (flux_hllc.F:344-391) on a
    two-stream grid: status, fallback count and every field as the oracle."""
    name, axis, seed, factor, status, nfb = case
    g, glob, _ = problems.two_stream(32, axis=axis, rank=3, seed=seed, ncolours=1)
    glob = dict(glob, RiemannSolver=4)
    g.dt = oracle.compute_time_step("port", g, glob, 0.4) * factor
    a, b = g.copy(), g.copy()
    assert oracle.solve_ppm_de("port", a, glob, cycle=axis) == (status, nfb)
    assert run(cpulib, b, glob, axis) == (status, nfb)
    if status == 1:
        assert all(r[1] for r in field_report(b, a))


@pytest.mark.parametrize("kind,solver", [("blast", 4), ("diffusion", 1), ("long", 1), ("long", 4)],
                         ids=["blast-hllc", "diffusion-hll", "long-hll", "long-hllc"])
def test_level_alternative_solvers(cpulib, kind, solver):
    """RiemannSolver HLL / HLLC on a batched level (every pencil group in
    the one-pencil workgroups, slice by slice with diffusion): fields and
    Courant steps equal to the per-grid oracle loop."""
    grids, glob = level_grids(kind, 3, seed=13)
    level_check(cpulib, grids, dict(glob, RiemannSolver=solver), 2)


# Work-item order and LDS initialisation.  The CPU build runs the work-items
# of a barrier interval one after another; on the GPU they run at once and
# LDS starts with whatever the previous workgroup left.  Here every __shared__
# array starts as NaN, a read of an LDS cell the workgroup has not written is
# counted (HIPCPU_LDS_CHECK), and each case runs with the work-items of every
# interval in forward, reverse and two shuffled orders: the results must be
# bitwise the same and no read may precede its write.
ORDER_CASES = ["1d_ok", "3d_x_hll_fails", "3d_y_ok", "3d_y_diffusion_hll_fails", "3d_z_stretched_ok"]


def _run_fields(lib, g, glob, cycle, fluxes=None):
    b = g.copy()
    fl = SubgridFluxes(b, fluxes) if fluxes else None
    st, n = run(lib, b, glob, cycle, fluxes=fl)
    return st, n, [f.copy() for f in b.fields], (fl.flat().copy() if fl else None)


@pytest.mark.parametrize("name", ORDER_CASES)
def test_work_item_order_and_lds_reads(cpulib, name, monkeypatch):
    cpulib.hipcpu_lds_violations.restype = C.c_ulonglong
    monkeypatch.setenv("HIPCPU_LDS_CHECK", "1")
    monkeypatch.setenv("HIPCPU_LDS_CONTINUE", "1")
    g, glob, cycle = fallback_case(name, ncolours=2)
    boxes = [((3, 15, 3), (8, 18, 7)), ((4, 20, 4), (6, 25, 6))] if name.startswith("3d_y") else None
    before = cpulib.hipcpu_lds_violations()
    out = {}
    for order in ("fwd", "rev", "shuf:1", "shuf:5"):
        monkeypatch.setenv("HIPCPU_ORDER", order)
        out[order] = _run_fields(cpulib, g, glob, cycle, boxes)
    assert cpulib.hipcpu_lds_violations() == before, "LDS read before write"
    st, n, fields, flux = out["fwd"]
    assert n > 0 and st == FALLBACK_CASES[name][-1]
    for order, (st2, n2, fields2, flux2) in out.items():
        assert (st2, n2) == (st, n), order
        assert all(bit_equal(x, y) for x, y in zip(fields, fields2)), order
        if flux is not None:
            assert bit_equal(flux, flux2), order


@pytest.mark.parametrize("name", ["diffusion", "turbulent", "ragged"])
def test_steps_lds_reads(cpulib, name, monkeypatch):
    """Plain steps (the main tile kernel, slice-serial diffusion launches)
    with every LDS read checked against the workgroup's own writes and the
    work-items run in reverse order."""
    cpulib.hipcpu_lds_violations.restype = C.c_ulonglong
    monkeypatch.setenv("HIPCPU_LDS_CHECK", "1")
    monkeypatch.setenv("HIPCPU_LDS_CONTINUE", "1")
    monkeypatch.setenv("HIPCPU_ORDER", "rev")
    before = cpulib.hipcpu_lds_violations()
    g, glob, _ = CASES[name]()
    fill_bc(g, "periodic")
    g.dt = oracle.compute_time_step("port", g, glob, 0.4)
    a, b = g.copy(), g.copy()
    assert oracle.solve_ppm_de("port", a, glob, cycle=1)[0] == 1
    assert run(cpulib, b, glob, 1) == (1, 0)
    assert all(r[1] for r in field_report(b, a))
    assert cpulib.hipcpu_lds_violations() == before, "LDS read before write"


GOLD_FLATTEN = sorted(__import__("glob").glob(__import__("os").path.join(
    __import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden",
    "step_*flatten*.npz")))


@pytest.mark.parametrize("path", GOLD_FLATTEN, ids=[p.rsplit("/", 1)[-1] for p in GOLD_FLATTEN])
def test_golden_flattening(cpulib, path):
    """The reference's own outputs with flattening types 2 and 3 on strong
    shocks (tests/golden), replayed bitwise by the library's source."""
    import test_golden as tg
    z, g, glob = tg.load_case(path)

    def step(g, s, fl):
        assert device.solve_ppm_de(g, glob, cycle=s, fluxes=fl, lib=cpulib) == 1

    tg.replay(z, g, glob, step, path.rsplit("/", 1)[-1])


@pytest.mark.parametrize("n", [24, 72])
def test_host_round_trip(cpulib, n):
    """The host-array transfers: the direct path (24^3) and the pinned
    staging ring (72^3: 13 fields of 3.3 MB, more than the ring's three
    buffers)."""
    from helpers import check_host_round_trip
    check_host_round_trip(cpulib, n)


def test_periodic_fill(cpulib):
    from helpers import check_periodic_fill
    check_periodic_fill(cpulib)
