"""Seeded synthetic initial conditions for the PPM-DE workloads.

Every generator returns (Grid, globals_dict, run_dict).  The states follow
the reference's conventions for a dual-energy ideal gas: specific gas energy
ge = p / ((gamma-1) rho), specific total energy E = ge + v^2/2
(Grid_HydroShockTubesInitializeGrid.C:38-49,75-78).
"""
import numpy as np

from .grid import Grid, periodic_fill

GAMMA = 5.0 / 3.0


def _set_thermo(g, rho, p, vx, vy, vz, gamma):
    ge = p / ((gamma - 1.0) * rho)
    g.d[...] = rho
    g.u[...] = vx
    g.v[...] = vy
    g.w[...] = vz
    g.e[...] = ge + 0.5 * (vx * vx + vy * vy + vz * vz)
    if g.dual_energy:
        g.ge[...] = ge


def sod_x(n=64, rank=3, gamma=GAMMA, dual_energy=True, flags=(1, 1, 1), cfl=0.4):
    """Sod shock tube along x on an n^rank periodic grid (x < 0.5: rho = 1,
    p = 1; else rho = 0.125, p = 0.1; v = 0).  flags = (steepening,
    flattening, diffusion)."""
    g = Grid([n] * rank, rank=rank, dual_energy=dual_energy, dx=1.0 / n)
    x = g.cell_centers(0)[None, None, :]
    left = x < 0.5
    shape = g.shape
    rho = np.where(left, 1.0, 0.125) * np.ones(shape)
    p = np.where(left, 1.0, 0.1) * np.ones(shape)
    z = np.zeros(shape)
    _set_thermo(g, rho, p, z, z, z, gamma)
    glob = dict(Gamma=gamma, PPMSteepeningParameter=flags[0], PPMFlatteningParameter=flags[1],
                PPMDiffusionParameter=flags[2], DualEnergyFormalism=int(dual_energy))
    return g, glob, dict(cfl=cfl, bc="periodic")


def gaussian_random_field(n, kmin, kmax, slope, rng, ncomp=1, sparse=False):
    """Real periodic GRF on an n^3 grid with power P(k) ~ k^slope on integer
    shells kmin <= |k| <= kmax (zero mean, unit variance per component).
    sparse: draw random numbers only for the modes inside the shells (a
    different seeded field, much cheaper at large n)."""
    kx = np.fft.fftfreq(n, d=1.0 / n)
    if sparse:
        kc = int(kmax) + 1
        lo = kx[np.abs(kx) <= kc]
        kr_x = np.fft.rfftfreq(n, d=1.0 / n)
        kz, ky, kxx = np.meshgrid(lo, lo, kr_x[kr_x <= kc], indexing="ij")
    else:
        kz, ky, kxx = np.meshgrid(kx, kx, np.fft.rfftfreq(n, d=1.0 / n), indexing="ij")
    k = np.sqrt(kxx * kxx + ky * ky + kz * kz)
    kr = np.rint(k)
    mask = (kr >= kmin) & (kr <= kmax)
    amp = np.zeros_like(k)
    amp[mask] = np.sqrt(k[mask] ** slope / (4.0 * np.pi * np.maximum(k[mask], 1.0) ** 2))
    out = []
    for _ in range(ncomp):
        ph = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
        if sparse:
            full = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
            iz = (kz % n).astype(np.intp)
            iy = (ky % n).astype(np.intp)
            full[iz, iy, kxx.astype(np.intp)] = ph * amp
            f = np.fft.irfftn(full, s=(n, n, n), axes=(0, 1, 2))
            del full
        else:
            f = np.fft.irfftn(ph * amp, s=(n, n, n), axes=(0, 1, 2))
        f -= f.mean()
        f /= f.std()
        out.append(f)
    return out


def turbulent(n=256, seed=2024, sigma_lnrho=0.5, mach=3.0, T=1.0, ncolours=3,
              gamma=GAMMA, flags=(1, 1, 0), cfl=0.4):
    """Uniform n^3 periodic grid: lognormal density (ln rho a k^-2 Gaussian
    random field on k = 1..16 with standard deviation sigma_lnrho, rho
    normalised to mean 1), velocity a k^-2 Gaussian random field on k = 1..16
    normalised to RMS Mach `mach` for c_s = sqrt(gamma T), p = rho T, and
    `ncolours` passive colours uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    g = Grid([n] * 3, ncolours=ncolours, dual_energy=True, dx=1.0 / n)
    lnrho = gaussian_random_field(n, 1, 16, -2.0, rng)[0] * sigma_lnrho
    rho = np.exp(lnrho)
    rho /= rho.mean()
    vel = gaussian_random_field(n, 1, 16, -2.0, rng, ncomp=3)
    cs = np.sqrt(gamma * T)
    vrms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    scale = mach * cs / vrms
    vel = [v * scale for v in vel]
    p = rho * T
    a = g.active_slice()
    full = lambda x: _embed(g, x)
    _set_thermo(g, full(rho), full(p), full(vel[0]), full(vel[1]), full(vel[2]), gamma)
    for c in g.colnum:
        g.fields[c][a] = rng.uniform(0.0, 1.0, size=(n, n, n))
    g.fill_periodic()
    glob = dict(Gamma=gamma, PPMSteepeningParameter=flags[0], PPMFlatteningParameter=flags[1],
                PPMDiffusionParameter=flags[2], DualEnergyFormalism=1)
    return g, glob, dict(cfl=cfl, bc="periodic")


def cosmological(n=512, seed=5, sigma_lnrho=2.0, mach=100.0, T=1.0, ncolours=6, g_frac=0.05,
                 gamma=GAMMA, flags=(1, 1, 0), cfl=0.4, log=None):
    """Comoving-box state: lognormal density (ln rho a k^-2 Gaussian random
    field on k = 1..16 with standard deviation sigma_lnrho, mean rho 1),
    k^-2 Gaussian random velocity at RMS Mach `mach` for c_s = sqrt(gamma T),
    p = rho T (so the thermal energy is ~1e-4 of the kinetic one and the
    dual-energy switches decide), `ncolours` species fractions uniform in
    [0, 1], and a static seeded 3-vector acceleration (a k^-2 Gaussian random
    field per component, RMS g_frac * v_rms^2 / L over a unit box).  Cell
    widths are comoving (1/n); the caller scales them by a."""
    rng = np.random.default_rng(seed)
    say = log or (lambda *a: None)
    g = Grid([n] * 3, ncolours=ncolours, dual_energy=True, gravity=True, dx=1.0 / n)
    a = g.active_slice()
    say("density")
    lnrho = gaussian_random_field(n, 1, 16, -2.0, rng, sparse=True)[0] * sigma_lnrho
    rho = np.exp(lnrho)
    del lnrho
    rho /= rho.mean()
    cs = np.sqrt(gamma * T)
    say("velocity")
    vel = gaussian_random_field(n, 1, 16, -2.0, rng, ncomp=3, sparse=True)
    vrms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    for v in vel:
        v *= mach * cs / vrms
    full = lambda x: _embed(g, x)  # noqa: E731
    _set_thermo(g, full(rho), full(rho * T), full(vel[0]), full(vel[1]), full(vel[2]), gamma)
    del vel, rho
    say("colours")
    for c in g.colnum:
        g.fields[c][a] = rng.uniform(0.0, 1.0, size=(n, n, n))
    say("acceleration")
    acc = gaussian_random_field(n, 1, 16, -2.0, rng, ncomp=3, sparse=True)
    arms = np.sqrt(np.mean(acc[0] ** 2 + acc[1] ** 2 + acc[2] ** 2))
    for d in range(3):
        g.accel[d][a] = acc[d] * (g_frac * (mach * cs) ** 2 / arms)
    del acc
    g.fill_periodic()
    for d in range(3):
        periodic_fill(g.accel[d], g.ghost, g.rank)
    glob = dict(Gamma=gamma, PPMSteepeningParameter=flags[0], PPMFlatteningParameter=flags[1],
                PPMDiffusionParameter=flags[2], DualEnergyFormalism=1)
    return g, glob, dict(cfl=cfl, bc="periodic")


def _embed(g, x):
    out = np.zeros(g.shape)
    out[g.active_slice()] = x
    periodic_tmp = out
    from .grid import periodic_fill
    periodic_fill(periodic_tmp, g.ghost, g.rank)
    return periodic_tmp


def amr_level(ngrids=512, seed=7, nmin=8, nmax=64, level_dx=1.0 / 256, gamma=GAMMA,
              ncolours=0, flags=(1, 1, 0), blast=1e4, radius=0.1):
    """One AMR level of `ngrids` uneven subgrids.  Each axis length is drawn
    log-uniform in [nmin, nmax]; each grid holds smooth sinusoidal density and
    velocity fields plus a Sedov-like blast (pressure x `blast` inside
    `radius` of a random centre inside the grid).  Ghost zones are filled
    from the same analytic fields."""
    rng = np.random.default_rng(seed)
    grids = []
    for _ in range(ngrids):
        ext = np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), size=3)).astype(int)
        ext = np.clip(ext, nmin, nmax)
        origin = rng.uniform(0.0, 1.0, size=3)
        g = Grid(list(ext), ncolours=ncolours, dual_energy=True, dx=level_dx,
                 left_edge=list(origin))
        x = g.cell_centers(0)[None, None, :]
        y = g.cell_centers(1)[None, :, None]
        z = g.cell_centers(2)[:, None, None]
        ph = rng.uniform(0, 2 * np.pi, size=4)
        k = 2 * np.pi * rng.integers(1, 4, size=4)
        rho = 1.0 + 0.3 * np.sin(k[0] * x + ph[0]) * np.sin(k[1] * y + ph[1]) * np.cos(k[2] * z + ph[2])
        vx = 0.5 * np.sin(k[1] * y + ph[3])
        vy = 0.5 * np.sin(k[2] * z + ph[0])
        vz = 0.5 * np.sin(k[0] * x + ph[1])
        size = ext * level_dx
        c = origin + rng.uniform(0.2, 0.8, size=3) * size
        r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
        p = np.ones(g.shape) * 1.0
        p = np.where(r < radius * size.max(), blast, 1.0) * np.ones(g.shape)
        shape = g.shape
        _set_thermo(g, rho * np.ones(shape), p, vx * np.ones(shape), vy * np.ones(shape),
                    vz * np.ones(shape), gamma)
        for cnum in g.colnum:
            g.fields[cnum][...] = rng.uniform(0.0, 1.0, size=shape)
        grids.append(g)
    glob = dict(Gamma=gamma, PPMSteepeningParameter=flags[0], PPMFlatteningParameter=flags[1],
                PPMDiffusionParameter=flags[2], DualEnergyFormalism=1)
    return grids, glob, dict(cfl=0.4, bc="analytic")


def random_smooth(active, seed=0, ncolours=0, dual_energy=True, gravity=False, rank=3,
                  amp=0.3, vamp=0.8, flags=(1, 1, 0), stretch=0.0):
    """Small seeded test state: smooth random density/pressure/velocity with
    moderate contrasts, used by the parity tests."""
    rng = np.random.default_rng(seed)
    g = Grid(active, rank=rank, ncolours=ncolours, dual_energy=dual_energy, gravity=gravity)
    if stretch:  # non-uniform cell widths (CellWidthTemp arrays)
        for d in range(rank):
            w = g.cellwidth[d]
            w *= 1.0 + stretch * np.sin(np.arange(w.size) * 0.7 + d)
    shape = g.shape
    x = g.cell_centers(0)[None, None, :]
    y = g.cell_centers(1)[None, :, None] if rank > 1 else np.zeros((1, 1, 1))
    z = g.cell_centers(2)[:, None, None] if rank > 2 else np.zeros((1, 1, 1))
    tp = 2 * np.pi

    def smooth():
        k = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, tp, size=3)
        return (np.sin(tp * k[0] * x + ph[0]) * np.cos(tp * k[1] * y + ph[1]) *
                np.sin(tp * k[2] * z + ph[2] + 0.3))

    rho = np.exp(amp * smooth() + 0.5 * amp * smooth()) * np.ones(shape)
    p = np.exp(amp * smooth()) * np.ones(shape)
    vx = vamp * smooth() * np.ones(shape)
    vy = (vamp * smooth() if rank > 1 else 0.0) * np.ones(shape)
    vz = (vamp * smooth() if rank > 2 else 0.0) * np.ones(shape)
    # add a few sharp features so limiters, flattening and shocks are exercised
    jump = (x > 0.5) & (y >= 0.0)
    rho = np.where(jump, rho * 0.3, rho)
    p = np.where(jump, p * 0.2, p)
    _set_thermo(g, rho, p, vx, vy, vz, GAMMA)
    for c in g.colnum:
        g.fields[c][...] = rng.uniform(0.0, 1.0, size=shape)
    if gravity:
        for a in range(3):
            g.accel[a][...] = 0.1 * smooth() * np.ones(shape)
    glob = dict(Gamma=GAMMA, PPMSteepeningParameter=flags[0], PPMFlatteningParameter=flags[1],
                PPMDiffusionParameter=flags[2], DualEnergyFormalism=int(dual_energy))
    return g, glob, dict(cfl=0.4, bc="periodic")


def two_stream(n=32, axis=0, rank=1, seed=0, noise=0.3, p0=1e-3, transverse=(6, 5), gamma=GAMMA,
               ncolours=0):
    """Cold gas streaming apart from the middle of axis `axis` (u = -1 for
    x < 0.5, +1 beyond, each times 1 + 0.2 N(0,1)), density lognormal with
    standard deviation `noise`, pressure p0.  The central rarefaction drains
    the middle cells; with a time step a few times the Courant one the
    two-shock fluxes empty one of them, which is the per-cell fallback case
    of flux_twoshock.F:401-456."""
    rng = np.random.default_rng(seed)
    act = [n] if rank == 1 else ([n] + list(transverse[:rank - 1]))
    if rank > 1:  # put the streaming axis at `axis`
        act = list(transverse[:rank - 1])
        act.insert(axis, n)
    g = Grid(act, rank=rank, ncolours=ncolours, dual_energy=True, dx=1.0 / n)
    x = [g.cell_centers(0)[None, None, :], g.cell_centers(1)[None, :, None] if rank > 1 else 0.0,
         g.cell_centers(2)[:, None, None] if rank > 2 else 0.0][axis] * np.ones(g.shape)
    rho = np.exp(noise * rng.standard_normal(g.shape))
    vel = np.where(x < 0.5, -1.0, 1.0) * (1 + 0.2 * rng.standard_normal(g.shape))
    p = p0 * np.ones(g.shape)
    z = np.zeros(g.shape)
    v = [z, z, z]
    v[axis] = vel
    _set_thermo(g, rho, p, v[0], v[1], v[2], gamma)
    for c in g.colnum:
        g.fields[c][...] = rng.uniform(0.0, 1.0, size=g.shape)
    g.fill_periodic()
    glob = dict(Gamma=gamma, PPMSteepeningParameter=1, PPMFlatteningParameter=1,
                PPMDiffusionParameter=0, DualEnergyFormalism=1)
    return g, glob, dict(cfl=0.4, bc="periodic")
