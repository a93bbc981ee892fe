"""Shared helpers of the tests (comparisons and run loops)."""
import os
import subprocess

import numpy as np

import oracle
from enzo_ppm import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CPU_LIB = os.path.join(HERE, "emu", "libenzo_ppm_cpu.so")
CSRC = os.path.join(ROOT, "enzo-dev_amd", "csrc")
HIP_UNITS = ["enzo_ppm.hip", "ppm_level.hip", "ppm_learn.hip"]


def cpu_library():
    """The library's own HIP source built for the CPU (tests/emu/hipcpu),
    rebuilt when a source is newer; loaded through the same C ABI."""
    src = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".hpp"))]
    if not os.path.exists(CPU_LIB) or os.path.getmtime(CPU_LIB) < max(map(os.path.getmtime, src)):
        subprocess.run(["g++", "-x", "c++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
                        "-Wno-attributes", "-U_FORTIFY_SOURCE", "-Wl,-Bsymbolic",
                        "-ffp-contract=off", "-fno-fast-math", "-I",
                        os.path.join(HERE, "emu", "hipcpu"), "-o", CPU_LIB]
                       + [os.path.join(CSRC, f) for f in HIP_UNITS], check=True)
    return abi.load_library(CPU_LIB)


def bit_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def field_report(g_test, g_ref):
    """[(field number, bitwise equal, max relative difference)]"""
    return [(k, bit_equal(x, y), max_rel(x, y)) for k, (x, y) in enumerate(zip(g_test.fields, g_ref.fields))]


def fill_bc(g, bc):
    if bc == "periodic":
        g.fill_periodic()
    elif bc == "outflow":
        g.fill_outflow()


def oracle_steps(kind, g, glob, nsteps, cfl, bc="periodic", cycle0=0, nthreads=8):
    """nsteps full steps on the oracle: ghost fill, Courant dt, SolvePPM_DE."""
    fb_total = 0
    for n in range(nsteps):
        fill_bc(g, bc)
        g.dt = oracle.compute_time_step(kind, g, glob, cfl)
        st, fb = oracle.solve_ppm_de(kind, g, glob, cycle=cycle0 + n, nthreads=nthreads)
        assert st == 1, f"oracle {kind} status {st}"
        fb_total += fb
    return fb_total


# Per-cell HLL fallback cases (flux_twoshock.F:401-456): the two-stream
# rarefaction of problems.two_stream with the time step set to `factor` times
# the Courant one (CFL 0.4), a factor bisected to the threshold where the
# first cell's two-shock fluxes empty it.  Status 1: HLL repairs it; -3: the
# HLL fluxes fail too (the reference stops).  Fields:
# (rank, axis, seed, factor, diffusion, stretch, expected status).
FALLBACK_CASES = {
    "1d_ok": (1, 0, 0, 3.163784086704254, 0, 0.0, 1),
    "1d_hll_fails": (1, 0, 2, 2.573289141058922, 0, 0.0, -3),
    "3d_x_ok": (3, 0, 1, 3.2654410749673843, 0, 0.0, 1),
    "3d_x_hll_fails": (3, 0, 0, 3.210948911011219, 0, 0.0, -3),
    "3d_y_ok": (3, 1, 1, 2.975198045372963, 0, 0.0, 1),
    "3d_z_ok": (3, 2, 5, 3.2506836354732513, 0, 0.0, 1),
    "3d_z_hll_fails": (3, 2, 5, 3.348204144537449, 0, 0.0, -3),
    "3d_x_stretched_ok": (3, 0, 1, 3.9790502339601517, 0, 0.2, 1),
    "3d_z_stretched_ok": (3, 2, 1, 3.3523254990577698, 0, 0.2, 1),
    "3d_y_diffusion_hll_fails": (3, 1, 1, 3.023725539445877, 1, 0.0, -3),
}


def evolved(case, steps, cfl=0.4):
    """(grid, globals, run) of `case` after `steps` periodic oracle steps:
    strong-shock states from smooth initial conditions, where the
    flattening types differ."""
    g, glob, run = case
    oracle_steps("port", g, glob, steps, cfl, "periodic")
    return g, glob, run


def sod_level(flatten, sizes=(16, 20, 12)):
    """A level of 3-D Sod tubes along x (shock, contact and rarefaction in
    every slice's first line, where the flattening of each slice is taken;
    diffusion on) with PPMFlatteningParameter `flatten`."""
    from enzo_ppm import problems
    base = [problems.sod_x(n) for n in sizes]
    return [b[0] for b in base], dict(base[0][1], PPMFlatteningParameter=flatten)


def check_host_round_trip(lib, n, seed=3):
    """hipPPMSetBaryon then hipPPMGetBaryon of an n^3 grid with gravity and
    colours (the pinned staging ring: fields of more than one 32 MiB chunk
    when n >= 160, more fields than ring buffers): every array comes back
    bit for bit, and an upload after a download (the ring's buffers reused
    in the other direction) too."""
    import numpy as np
    from enzo_ppm import device, problems
    g, glob, _ = problems.random_smooth([n, n, n], seed=seed, ncolours=2, gravity=True)
    rng = np.random.default_rng(seed)
    for f in g.fields:
        f[...] = rng.standard_normal(f.shape)
    want = [f.copy() for f in g.fields]
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        for f in g.fields:
            f[...] = 0.0
        assert dg.download() == 1
        for a, b in zip(g.fields, want):
            assert bit_equal(a, b)
        for f in g.fields:
            f *= 2.0
        dg.upload()
        for f in g.fields:
            f[...] = 0.0
        assert dg.download() == 1
        for a, b in zip(g.fields, want):
            assert bit_equal(a, 2.0 * b)
    finally:
        dg.close()


def fallback_case(name, ncolours=1):
    """(grid, globals, cycle) of a fallback case with grid.dt set; the first
    sweep of `cycle` runs along the streaming axis."""
    from enzo_ppm import problems
    rank, axis, seed, factor, diff, stretch, _ = FALLBACK_CASES[name]
    g, glob, _ = problems.two_stream(32, axis=axis, rank=rank, seed=seed, ncolours=ncolours)
    glob = dict(glob, PPMDiffusionParameter=diff)
    if stretch:
        for d in range(3):
            w = g.cellwidth[d]
            w *= 1.0 + stretch * np.sin(np.arange(w.size) * 0.7 + d)
    g.dt = oracle.compute_time_step("port", g, glob, 0.4) * factor
    return g, glob, axis


def fallback_differences(grid, fields_ref, nfb):
    """The oracle's fields against the reference's after a fallback step:
    every field bitwise equal except the gas energy in at most `nfb` cells
    (the cell after each fallback cell, whose gas-energy source the
    reference's one-cell flux_hll call reads from unset locals,
    flux_hll.F:304-307).  Returns the number of such cells."""
    ge = 5 if grid.dual_energy else None
    n = 0
    for k, (x, y) in enumerate(zip(grid.fields, fields_ref)):
        d = int(np.count_nonzero(x.view(np.int64) != np.asarray(y).view(np.int64)))
        if k == ge:
            n = d
        else:
            assert d == 0, (k, d)
    assert n <= nfb, (n, nfb)
    return n


def check_tune_times(lv, choice):
    """The tuner kept, per pencil-length class, the shape it timed fastest,
    and the faster of the two launch forms (hipPPMLevelTuneTimes)."""
    import numpy as np
    times, (plain, pers) = lv.tune_times()
    for c in range(4):
        row = times[c]
        if np.all(np.isnan(row)):
            continue
        assert np.all(row[~np.isnan(row)] > 0)
        assert choice[c] == int(np.nanargmin(row)), (c, row, choice)
    assert plain > 0 and pers > 0
    assert lv.persistent() == (pers < plain)
    return times, (plain, pers)


def level_reference(grids, glob, steps, cfl=0.4, dt=None):
    """The oracle's per-grid loop for a level: each step every grid takes the
    level's dt (the minimum over the grids of the Courant step, or `dt`) and
    one SolvePPM_DE call; ghosts are left as they are.  Returns the dts."""
    dts = []
    for s in range(steps):
        d = dt if dt is not None else min(oracle.compute_time_step("port", g, glob, cfl) for g in grids)
        for g in grids:
            g.dt = d
            st, _ = oracle.solve_ppm_de("port", g, glob, cycle=s)
            assert st == 1
        dts.append(d)
    return dts


def level_grids(kind, n=4, seed=7):
    """Small seeded levels for the batched-update tests."""
    from enzo_ppm import problems
    if kind == "blast":
        return problems.amr_level(n, seed=seed, nmin=8, nmax=20, ncolours=1)[:2]
    if kind == "gravity":
        out = []
        for i in range(n):
            g, glob, _ = problems.random_smooth([8 + 2 * i, 10, 9 + i], seed=seed + i, ncolours=1,
                                                gravity=True)
            g.fill_periodic()
            out.append(g)
        return out, glob
    if kind == "stretched":
        out = []
        for i in range(n):
            g, glob, _ = problems.random_smooth([12, 9 + i, 10], seed=seed + i, ncolours=2,
                                                stretch=0.15)
            g.fill_periodic()
            out.append(g)
        return out, glob
    if kind == "long":  # ~200- and ~400-cell pencils: 256- and 512-wide one-pencil workgroups
        out = []
        for i, act in enumerate(([200, 4, 3], [5, 400, 2], [6, 3, 5])[:n]):
            g, glob, _ = problems.random_smooth(act, seed=seed + i, ncolours=1)
            g.fill_periodic()
            out.append(g)
        return out, glob
    if kind == "diffusion":
        out = []
        for i in range(n):
            g, glob, _ = problems.random_smooth([9 + i, 8, 10], seed=seed + i, ncolours=1,
                                                flags=(1, 1, 1))
            g.fill_periodic()
            out.append(g)
        return out, glob
    raise ValueError(kind)


def oracle_comoving_steps(kind, g, glob, model, t, nsteps, cfl, cycle0=0, nthreads=8):
    """`nsteps` comoving steps on the oracle in Enzo's order: periodic fill,
    Courant step with a(t) on the unscaled widths, SolvePPM_DE with the
    widths scaled by a(t + dt/2), then the expansion terms with the pressure
    at the step's mid-time.  Returns (t, [dt per step])."""
    dts = []
    for n in range(nsteps):
        a_now, _ = model(t)
        g.fill_periodic()
        dt = oracle.compute_time_step(kind, g, glob, cfl, aye=a_now)
        a_half, _ = model(t + 0.5 * dt)
        old = g.copy()
        cw = g.cellwidth
        g.cellwidth = [a_half * c if d < g.rank else c for d, c in enumerate(cw)]
        g.dt = dt
        try:
            st, _ = oracle.solve_ppm_de(kind, g, glob, cycle=cycle0 + n, nthreads=nthreads)
        finally:
            g.cellwidth = cw
        assert st == 1, f"oracle {kind} status {st}"
        t_new = t + dt
        a_mid, dadt = model(0.5 * (t_new + t))
        coef = (0.5 * (t_new + t) - t) / (t_new - t)
        oracle.comoving_expansion_terms(kind, g, old, glob, coef, dt * dadt / a_mid)
        t = t_new
        dts.append(dt)
    return t, dts


def oracle_comoving_level(grids, glob, model, t, nsteps, cfl, boundary=True):
    """The oracle's per-grid loop for a comoving level: periodic fill of
    every grid (when `boundary`), the level's Courant step (minimum over the
    grids, with a(t)), then per grid SolvePPM_DE on a(t+dt/2)-scaled widths
    and the expansion terms.  Returns (t, [dt per step])."""
    dts = []
    for n in range(nsteps):
        a_now, _ = model(t)
        if boundary:
            for g in grids:
                g.fill_periodic()
        dt = min(oracle.compute_time_step("port", g, glob, cfl, aye=a_now) for g in grids)
        a_half, _ = model(t + 0.5 * dt)
        t_new = t + dt
        a_mid, dadt = model(0.5 * (t_new + t))
        coef = (0.5 * (t_new + t) - t) / (t_new - t)
        for g in grids:
            old = g.copy()
            cw = g.cellwidth
            g.cellwidth = [a_half * c if d < g.rank else c for d, c in enumerate(cw)]
            g.dt = dt
            try:
                st, _ = oracle.solve_ppm_de("port", g, glob, cycle=n)
            finally:
                g.cellwidth = cw
            assert st == 1
            oracle.comoving_expansion_terms("port", g, old, glob, coef, dt * dadt / a_mid)
        t = t_new
        dts.append(dt)
    return t, dts


def level_subgrid_boxes(g):
    """Subgrid boxes of one level grid, in its own index space: an interior
    box, the whole active region, and a one-cell-thick slab on the first
    axis (faces on the grid's outer edges and inside it)."""
    s, e = list(g.start), list(g.end)
    inner = ([min(s[d] + 2, e[d]) for d in range(3)], [max(e[d] - 2, s[d] + 2) for d in range(3)])
    inner = (inner[0], [max(inner[0][d], inner[1][d]) for d in range(3)])
    mid = (s[0] + e[0]) // 2
    slab = ([mid] + s[1:], [mid] + e[1:])
    return [tuple(map(tuple, inner)), (tuple(s), tuple(e)), tuple(map(tuple, slab))]


def level_flux_check(lib, kind, steps=2, graph=True, capture=(0, 2), solver=None):
    """Subgrid face fluxes of a batched level: the grids in `capture` record
    the faces of level_subgrid_boxes; after `steps` steps the fields of every
    grid and the captured fluxes equal the oracle's per-grid SolvePPM_DE
    calls with the same SubgridFluxes, bitwise."""
    from enzo_ppm.grid import SubgridFluxes
    from enzo_ppm.level import Level
    grids, glob = level_grids(kind, 3)
    if solver is not None:
        glob = dict(glob, RiemannSolver=solver)
    ref = [g.copy() for g in grids]
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    try:
        lv.use_graph(graph)
        capture = [c for c in capture if c < len(grids)]
        fr = {c: SubgridFluxes(ref[c], level_subgrid_boxes(ref[c])) for c in capture}
        fl = {c: SubgridFluxes(lv.grids[c], level_subgrid_boxes(lv.grids[c])) for c in capture}
        for c in capture:
            lv.set_subgrid_fluxes(c, fl[c])
        for s in range(steps):
            d = min(oracle.compute_time_step("port", g, glob, 0.4) for g in ref)
            for i, g in enumerate(ref):
                g.dt = d
                st, _ = oracle.solve_ppm_de("port", g, glob, cycle=s, fluxes=fr.get(i))
                assert st == 1
            lv.step(s, 0.0, 0.4)
            st, dtg, _ = lv.status()
            assert st == 1 and dtg == d, (st, dtg, d)
        for c in capture:
            lv.save_subgrid_fluxes(c, fl[c])
            assert bit_equal(fr[c].flat(), fl[c].flat()), c
            assert np.abs(fr[c].flat()).max() > 0.0
        lv.download()
        for a, b in zip(lv.grids, ref):
            assert all(r[1] for r in field_report(a, b))
        # capture off: the level steps on, and the grid has no faces to save
        lv.set_subgrid_fluxes(capture[0], None)
        lv.step(steps, 0.0, 0.4)
        assert lv.status()[0] == 1
        try:
            lv.save_subgrid_fluxes(capture[0], fl[capture[0]])
        except RuntimeError:
            pass
        else:
            raise AssertionError("save with a stale face list was accepted")
    finally:
        lv.close()


def noalloc_solve(lib, g, glob, cycle, field_addrs, accel_addrs=None, min_support=0.0):
    """hipPPMSolvePPM_DE_NoAlloc on grid `g`'s description with the fields at
    the given addresses (device memory on the GPU; on the CPU build, host
    arrays); returns (status, fallback cells)."""
    import ctypes as C
    from enzo_ppm.grid import make_globals
    gs, _ = make_globals(**glob)
    cg, keep = g.ctypes_grid()
    bf = (abi.PF * abi.MAX_NUMBER_OF_BARYON_FIELDS)()
    for i, a in enumerate(field_addrs):
        bf[i] = C.cast(C.c_void_p(a), abi.PF)
    cg.BaryonField = C.cast(bf, C.POINTER(abi.PF))
    acc = None
    if accel_addrs is not None:
        acc = (abi.PF * 3)()
        for i, a in enumerate(accel_addrs):
            acc[i] = C.cast(C.c_void_p(a), abi.PF)
        cg.AccelerationField = C.cast(acc, C.POINTER(abi.PF))
    cw = g.ctypes_cellwidth()
    colnum = abi.iarr(g.colnum, max(1, len(g.colnum)))
    nfb = abi.LI(0)
    st = lib.hipPPMSolvePPM_DE_NoAlloc(C.byref(cg), C.byref(gs), cycle, cw, int(g.gravity),
                                       g.ncolours, colnum, min_support, None, C.byref(nfb))
    del keep, bf, acc
    return st, nfb.value


# cases of the entry point that allocates nothing: two calls in a row, the
# second on a smaller grid (per-cell fallback in the first; colours,
# gravity, stretched widths and diffusion in the second)
def noalloc_cases():
    from enzo_ppm import problems
    g1, glob1, cycle1 = fallback_case("3d_x_ok", ncolours=2)
    g2, glob2, _ = problems.random_smooth([11, 9, 12], seed=43, ncolours=1, gravity=True,
                                          stretch=0.1, flags=(1, 1, 1))
    g2.fill_periodic()
    g2.dt = oracle.compute_time_step("port", g2, glob2, 0.4)
    return [(g1, glob1, cycle1), (g2, glob2, 2)]


class DeviceArrays:
    """Device copies of float64 numpy arrays, made through the HIP runtime
    the library itself links (libamdhip64.so.7), for tests that hand the
    library device pointers."""

    def __init__(self, arrays):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs, self.sizes = [], []
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
            self.ptrs.append(p)
            self.sizes.append(a.nbytes)
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def addrs(self):
        return [p.value for p in self.ptrs]

    def read(self, k, out):
        assert out.flags["C_CONTIGUOUS"] and out.nbytes == self.sizes[k]
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptrs[k], out.nbytes, 2) == 0  # to host

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _carried(g):
    """Indices of the fields the device state holds: all but the velocity
    components beyond the grid's rank (fields d, e, vx, vy, vz, ge, colours)."""
    return [k for k in range(len(g.fields)) if not (2 + g.rank <= k <= 4)]


def _scramble_ghosts(g, seed):
    """Ghost cells of every field the device holds set to distinct seeded
    values (so that a fill that misses or misplaces a cell shows)."""
    rng = np.random.default_rng(seed)
    sl = g.active_slice()
    for k in _carried(g):
        f = g.fields[k]
        keep = f[sl].copy()
        f[...] = rng.uniform(-1e3, -1.0, f.shape)
        f[sl] = keep


def check_periodic_fill(lib):
    """Device periodic ghost fill, per grid and on a level of ragged grids,
    against the host fill (ExternalBoundary periodic case), every cell of
    every field bitwise; a grid whose active extent along an axis is below
    its ghost count is refused."""
    import pytest

    from enzo_ppm import device, problems
    from enzo_ppm.level import Level
    cases = [problems.random_smooth([7, 5, 4], seed=31, ncolours=2),
             problems.random_smooth([9, 6, 1], seed=32, ncolours=1, rank=2),
             problems.random_smooth([12, 1, 1], seed=33, rank=1)]
    for k, (g, glob, _) in enumerate(cases):
        _scramble_ghosts(g, 40 + k)
        ref = g.copy()
        ref.fill_periodic()
        dg = device.DeviceGrid(g, glob, lib=lib)
        try:
            dg.periodic_boundary()
            assert dg.download() == abi.SUCCESS
        finally:
            dg.close()
        assert all(bit_equal(g.fields[i], ref.fields[i]) for i in _carried(g)), k
    grids = [problems.random_smooth(a, seed=34 + i, ncolours=1)[0]
             for i, a in enumerate(([6, 4, 5], [3, 8, 4], [5, 5, 9]))]
    glob = problems.random_smooth([6, 4, 5], seed=34, ncolours=1)[1]
    for i, g in enumerate(grids):
        _scramble_ghosts(g, 50 + i)
    refs = [g.copy() for g in grids]
    for r in refs:
        r.fill_periodic()
    lv = Level(grids, glob, lib=lib)
    try:
        lv.periodic_boundary()
        lv.download()
    finally:
        lv.close()
    for i, (g, r) in enumerate(zip(grids, refs)):
        assert all(bit_equal(x, y) for x, y in zip(g.fields, r.fields)), i
    g, glob, _ = problems.random_smooth([2, 5, 4], seed=35)
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        with pytest.raises(RuntimeError):
            dg.periodic_boundary()
    finally:
        dg.close()
    lv = Level([g], glob, lib=lib)
    try:
        with pytest.raises(RuntimeError):
            lv.periodic_boundary()
    finally:
        lv.close()
