"""1-D shock tubes of the reference's answer tests
(run/Hydro/Hydro-1D/*/{*.enzo,*__test_*.py}): 100 cells on [0, 1], outflow
boundaries, PPM-DE with the default PPM options, gamma 1.4, CFL 0.9, run to
StopTime with the last step clipped to it; compared with the exact Riemann
solutions shipped beside those tests (copied under tests/golden as data)."""
import numpy as np

from .problems import _set_thermo
from .grid import Grid

TESTS = {
    # name: (x0, (rhoL, uL, pL), (rhoR, uR, pR), t_stop, exact file, fields, windows, rtol, atol)
    "sod": (0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.25, "SodShockTube_t0.25_exact.txt",
            ("density", "specific_thermal_energy"), ((0.25, 0.4), (0.85, 0.9)), 1e-2, 1e-7),
    "toro1": (0.3, (1.0, 0.75, 1.0), (0.125, 0.0, 0.1), 0.20, "Toro-1-ShockTube_t0.2_exact.txt",
              ("density", "velocity_x", "pressure"), ((0.0, 0.25), (0.65, 0.7)), 5e-2, 1e-7),
    "toro2": (0.5, (1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.15, "Toro-2-ShockTube_t0.15_exact.txt",
              ("density", "pressure"), ((0.2, 0.3), (0.7, 0.8)), 1e-1, 1e-7),
    "toro3": (0.5, (1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 0.012, "Toro-3-ShockTube_t0.012_exact.txt",
              ("density", "velocity_x", "pressure", "specific_thermal_energy"),
              ((0.1, 0.25), (0.6, 0.7)), 1e-1, 1e-7),
    "toro4": (0.4, (5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 0.035,
              "Toro-4-ShockTube_t0.035_exact.txt",
              ("density", "velocity_x", "pressure", "specific_thermal_energy"),
              ((0.5, 0.6), (0.78, 0.8)), 1e-1, 1e-7),
    "toro5": (0.8, (1.0, -19.59745, 1000.0), (1.0, -19.59745, 0.01), 0.012,
              "Toro-5-ShockTube_t0.012_exact.txt",
              ("density", "pressure", "specific_thermal_energy"), ((0.25, 0.4),), 1e-2, 1e-7),
    "toro6": (0.5, (1.4, 0.0, 1.0), (1.0, 0.0, 1.0), 2.0, "Toro-6-ShockTube_t2.0_exact.txt",
              ("density", "velocity_x", "pressure", "specific_thermal_energy"), ((0.0, 1.0),),
              1e-6, 1e-7),
    "toro7": (0.5, (1.4, 0.1, 1.0), (1.0, 0.1, 1.0), 2.0, "Toro-7-ShockTube_t2.0_exact.txt",
              ("density", "velocity_x", "pressure", "specific_thermal_energy"),
              ((0.0, 0.6), (0.8, 1.0)), 1e-3, 1e-7),
}

GLOBALS = dict(Gamma=1.4, DualEnergyFormalism=0)  # Enzo defaults otherwise
CFL = 0.9


def initial(name, n=100):
    x0, L, R, *_ = TESTS[name]
    g = Grid([n], rank=1, dual_energy=False, dx=1.0 / n)
    x = g.cell_centers(0)[None, None, :]
    left = x < x0
    rho = np.where(left, L[0], R[0]) * np.ones(g.shape)
    u = np.where(left, L[1], R[1]) * np.ones(g.shape)
    p = np.where(left, L[2], R[2]) * np.ones(g.shape)
    z = np.zeros(g.shape)
    _set_thermo(g, rho, p, u, z, z, 1.4)
    return g


def run(name, step_fn, dt_fn, n=100):
    """Integrate to StopTime: step_fn(grid, cycle) advances grid.dt."""
    g = initial(name, n)
    t_stop = TESTS[name][3]
    t, cycle = 0.0, 0
    while t < t_stop * (1 - 1e-12):
        g.fill_outflow()
        dt = dt_fn(g)
        dt = min(dt, t_stop - t)
        g.dt = dt
        step_fn(g, cycle)
        t += dt
        cycle += 1
    g.fill_outflow()
    return g, cycle


def load_exact(path):
    a = np.loadtxt(path, comments="#")
    return dict(pos=a[:, 0], density=a[:, 1], velocity_x=a[:, 2], pressure=a[:, 3],
                specific_thermal_energy=a[:, 4])


def compare(name, g, exact_path):
    """Returns [(field, window, max violation ratio)]; <= 1 passes the
    reference's assert_allclose(rtol, atol) on that window."""
    *_, fields, windows, rtol, atol = TESTS[name][:]
    ex = load_exact(exact_path)
    sl = g.active_slice()
    x = g.cell_centers(0)[g.ghost:g.ghost + g.active[0]]
    rho = g.d[sl].ravel()
    u = g.u[sl].ravel()
    eth = g.e[sl].ravel() - 0.5 * u * u
    vals = dict(density=rho, velocity_x=u, specific_thermal_energy=eth,
                pressure=(1.4 - 1.0) * rho * eth)
    out = []
    for f in fields:
        for (lo, hi) in windows:
            m = (x >= lo) & (x <= hi)
            want = np.interp(x[m], ex["pos"], ex[f])
            viol = np.abs(vals[f][m] - want) / (atol + rtol * np.abs(want))
            out.append((f, (lo, hi), float(viol.max())))
    return out
