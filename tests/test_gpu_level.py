"""Batched AMR-level update on the GPU (hipPPMLevel*) against the oracle's
per-grid loop: every grid bitwise equal to its own SolvePPM_DE with the
level's dt, the level dt equal to the minimum Courant step, fallback counts
summed over the grids."""
import pytest

import oracle
from enzo_ppm import abi, problems
from enzo_ppm.level import Level
from helpers import (bit_equal, check_tune_times, fallback_case, field_report, level_grids,
                     level_reference, sod_level)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


def level_check(lib, grids, glob, steps, shapes=None, graph=True, persistent=False):
    ref = [g.copy() for g in grids]
    dts = level_reference(ref, glob, steps)
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
            lv.step(s, 0.0, 0.4)
            st, dtg, nfb = lv.status()
            assert st == 1 and dtg == dts[s], (st, dtg, dts[s])
        lv.download()
        for a, b in zip(lv.grids, ref):
            rep = field_report(a, b)
            assert all(r[1] for r in rep), (a.active, rep)
    finally:
        lv.close()


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_level_blast_shapes(lib, shape):
    grids, glob = level_grids("blast", 4)
    level_check(lib, grids, glob, 3, shapes=shape)


@pytest.mark.parametrize("iflatten", [2, 3])
def test_level_flattening_types(lib, iflatten):
    """PPMFlatteningParameter 2 and 3 (calcdiss.F:447-566) on the batched
    level: a level of Sod tubes (strong shocks in every slice's first line,
    where each slice's flattening is taken; diffusion on), every grid
    bitwise as the oracle's, whose result differs from flattening type 1's."""
    grids, glob = sod_level(iflatten)
    level_check(lib, grids, glob, 3)
    a, b = [g.copy() for g in grids], [g.copy() for g in grids]
    level_reference(a, glob, 3)
    level_reference(b, dict(glob, PPMFlatteningParameter=1), 3)
    assert any(not bit_equal(x, y) for ga, gb in zip(a, b) for x, y in zip(ga.fields, gb.fields))


def test_level_direct_launches(lib):
    grids, glob = level_grids("blast", 3, seed=9)
    level_check(lib, grids, glob, 2, graph=False)


@pytest.mark.parametrize("kind", ["gravity", "stretched", "diffusion"])
def test_level_kinds(lib, kind):
    grids, glob = level_grids(kind, 3)
    level_check(lib, grids, glob, 2)


@pytest.mark.parametrize("kind", ["blast", "diffusion"])
@pytest.mark.parametrize("graph", [True, False])
def test_level_persistent(lib, kind, graph):
    """Persistent sweep launches (workgroups pulling work items from per-XCD
    counters, hipPPMLevelPersistent): every grid bitwise as the oracle's
    per-grid loop, over steps that reuse the counters."""
    grids, glob = level_grids(kind, 4)
    level_check(lib, grids, glob, 3, persistent=True, graph=graph)


@pytest.mark.parametrize("persistent", [False, True])
def test_level_many_grids(lib, persistent):
    """48 grids, axes log-uniform in [8, 64]: the AMR-level workload shape
    (with persistent launches: more work items than workgroups)."""
    grids, glob, _ = problems.amr_level(48, seed=21, nmin=8, nmax=64, ncolours=0)
    level_check(lib, grids, glob, 2, persistent=persistent)


def test_level_fallback(lib):
    grids = []
    for name in ("3d_x_ok", "3d_x_stretched_ok"):
        g, glob, _ = fallback_case(name)
        grids.append(g)
    dt = min(g.dt for g in grids)
    ref = [g.copy() for g in grids]
    nfb = 0
    for g in ref:
        g.dt = dt
        st, n = oracle.solve_ppm_de("port", g, glob, cycle=0)
        assert st == 1
        nfb += n
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        lv.step(0, dt, 0.4)
        st, _, n = lv.status()
        assert (st, n) == (1, nfb)
        lv.download()
        for a, b in zip(lv.grids, ref):
            assert all(r[1] for r in field_report(a, b))
    finally:
        lv.close()


def test_level_tune(lib):
    grids, glob, _ = problems.amr_level(24, seed=5, nmin=8, nmax=64, ncolours=0)
    dt = min(oracle.compute_time_step("port", g, glob, 0.4) for g in grids)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        choice = lv.tune(0, dt, repeats=2)
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


@pytest.mark.parametrize("parts", [(2, 2, 2), (4, 1, 2)])
def test_level_sibling_copies_match_oracle(lib, parts):
    import numpy as np
    import boundary_oracle as bo
    g, glob, _ = problems.turbulent(24, seed=3, ncolours=2)
    pieces, starts = bo.split_domain(g, parts)
    rng = np.random.default_rng(2)
    for h in pieces:
        sl = h.active_slice()
        for f in h.fields:
            act = f[sl].copy()
            f[...] = rng.normal(size=f.shape)
            f[sl] = act
    ref = [h.copy() for h in pieces]
    bo.copy_zones(ref, starts, g.active, periodic=True)
    lv = Level([h.copy() for h in pieces], glob, lib=lib)
    try:
        lv.set_siblings(starts, g.active, periodic=True)
        lv.copy_zones()
        lv.download()
    finally:
        lv.close()
    for a, b in zip(lv.grids, ref):
        assert all(r[1] for r in field_report(a, b))


def test_level_decomposed_box_equals_whole(lib):
    """A periodic 48^3 box as 8 sibling grids of 24^3 with device ghost
    copies before every level step equals the undivided box bitwise (dual
    energy and flattening off: with them the reference's result depends on
    where grid edges are)."""
    import boundary_oracle as bo
    from helpers import bit_equal
    g, glob, _ = problems.turbulent(48, seed=6, ncolours=1, flags=(1, 0, 0))
    glob = dict(glob, DualEnergyFormalism=0)
    pieces, starts = bo.split_domain(g, (2, 2, 2))
    whole = g.copy()
    lv = Level(pieces, glob, lib=lib)
    try:
        lv.set_siblings(starts, g.active, periodic=True)
        for s in range(3):
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
        w = tuple(slice(st0[d] + 2 * gh, st0[d] + 2 * gh + h.active[d]) for d in (2, 1, 0))
        sl = h.active_slice()
        for x, y in zip(h.fields, whole.fields):
            assert bit_equal(x[sl], y[w])


@pytest.mark.parametrize("solver", [1, 4], ids=["hll", "hllc"])
@pytest.mark.parametrize("kind", ["blast", "diffusion", "gravity", "long"])
@pytest.mark.parametrize("graph", [True, False])
def test_level_alternative_solvers(lib, kind, solver, graph):
    """RiemannSolver HLL / HLLC on a batched level: every pencil group in
    the one-pencil workgroups (slice by slice with diffusion), graph replay
    or direct launches; fields and Courant steps equal to the per-grid
    oracle loop."""
    grids, glob = level_grids(kind, 4, seed=13)
    level_check(lib, grids, dict(glob, RiemannSolver=solver), 3, graph=graph)


@pytest.mark.parametrize("solver", [5, 1, 4], ids=["twoshock", "hll", "hllc"])
@pytest.mark.parametrize("kind", ["blast", "diffusion", "gravity", "stretched"])
@pytest.mark.parametrize("graph", [True, False])
def test_level_subgrid_fluxes(lib, kind, solver, graph):
    """Subgrid face fluxes recorded by the batched level (captured graph or
    direct launches) equal the oracle's per-grid SolvePPM_DE with the same
    SubgridFluxes, bitwise, and so do the fields."""
    from helpers import level_flux_check
    print(f"level fluxes {kind} solver={solver} graph={graph}", flush=True)
    level_flux_check(lib, kind, graph=graph, solver=solver)
