"""Shared helpers of the tests (comparisons and run loops)."""
import numpy as np

import oracle


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
