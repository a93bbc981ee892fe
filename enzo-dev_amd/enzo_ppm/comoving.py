"""Comoving-coordinate steps on the device state, in the order Enzo takes
them (EvolveLevel.C -> grid::ComputeTimeStep, grid::SolveHydroEquations,
grid::ComovingExpansionTerms):

1. Courant step with a(Time) (Grid_ComputeTimeStep.C:103-106);
2. sweeps with the cell widths scaled by a(Time + dt/2)
   (Grid_SolveHydroEquations.C:386-408);
3. Time += dt, then the expansion terms with the pressure at
   0.5 (Time + OldTime) and Coefficient = dt (da/dt) / a there
   (Grid_ComovingExpansionTerms.C:55-130).

The expansion history a(t) is the caller's (Enzo's CosmologyComputeExpansion
Factor is outside the hydro path); `EdS` is a matter-only model for tests and
benchmarks.  Enzo's extra time-step limit on the expansion rate
(CosmologyComputeExpansionTimestep) is the caller's too.
"""


class EdS:
    """Einstein-de Sitter expansion a(t) = a0 (t / t0)^(2/3)."""

    def __init__(self, a0=0.1, t0=1.0):
        self.a0, self.t0 = float(a0), float(t0)

    def __call__(self, t):
        """(a, da/dt) at time t."""
        a = self.a0 * (t / self.t0) ** (2.0 / 3.0)
        return a, 2.0 / 3.0 * a / t


def device_step(dg, model, t, cycle, cfl, boundary=True):
    """One comoving step of a DeviceGrid with comoving handling on.  Returns
    (new time, dt)."""
    a_now, _ = model(t)
    dg.set_expansion_factors(a_now, dg.a_sweep)
    if boundary:
        dg.periodic_boundary()
    dt = dg.pressure_and_dt(cfl)
    a_half, _ = model(t + 0.5 * dt)
    dg.set_expansion_factors(a_now, a_half)
    dg.step(cycle, dt)
    t_new = t + dt
    a_mid, dadt = model(0.5 * (t_new + t))
    coef = (0.5 * (t_new + t) - t) / (t_new - t)
    dg.comoving_expansion_terms(coef, dt * dadt / a_mid)
    return t_new, dt


def level_step(lv, model, t, cycle, cfl, boundary=True):
    """One comoving step of a Level with comoving handling on (every grid
    takes the level's Courant step).  Returns (new time, dt)."""
    a_now, _ = model(t)
    lv.set_expansion_factors(a_now, lv.a_sweep)
    if boundary:
        lv.periodic_boundary()
    dt = lv.time_step(cfl)
    a_half, _ = model(t + 0.5 * dt)
    lv.set_expansion_factors(a_now, a_half)
    lv.step(cycle, dt, cfl)
    t_new = t + dt
    a_mid, dadt = model(0.5 * (t_new + t))
    coef = (0.5 * (t_new + t) - t) / (t_new - t)
    lv.comoving_expansion_terms(coef, dt * dadt / a_mid)
    return t_new, dt
