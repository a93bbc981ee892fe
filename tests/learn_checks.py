"""Checks of the learned two-shock start shared by the CPU-library tests and
the GPU tests (same C ABI, different library and sizes)."""
import numpy as np

import learn_oracle as lo
import oracle
from enzo_ppm import device, learn, problems
from helpers import field_report, oracle_steps

GAMMA = 5.0 / 3.0
PMIN = 1e-20   # MinimumPressure of the problem generators' globals


def harvest(lib, g, glob, steps=1, keep=1, cap=1 << 22, cfl=0.4):
    """Records and the interfaces selected of `steps` harvested steps."""
    return learn.harvest(g, glob, steps=steps, keep=keep, cap=cap, cfl=cfl, lib=lib)


def swept_interfaces(g):
    """Interfaces one full step solves: per axis (i2 - i1 + 2) per pencil,
    every transverse pencil (ghosts included) swept."""
    n = 0
    for d in range(g.rank):
        pencils = int(np.prod([g.dims[k] for k in range(3) if k != d]))
        n += (g.active[d] + 1) * pencils
    return n


def check_harvest_pstar(records):
    """The harvested p* is the reference twoshock.F's on the harvested
    inputs, bit for bit (and the oracle port's)."""
    kinds = ["port"] + (["ref"] if oracle.available("ref") else [])
    for kind in kinds:
        pbar, _ = lo.twoshock(kind, records, GAMMA, PMIN)
        assert np.array_equal(pbar, records[:, 6]), kind


def check_gradient(lib, records, seed=1):
    rng = np.random.default_rng(seed)
    w = learn.init_weights(seed)
    ok = lo.learnable(records, GAMMA)
    X, Y = lo.features(records[ok], GAMMA, fp32=True)   # the kernels' input rounding
    w[lo.OFF_MEAN:lo.OFF_INV_STD] = X.mean(0)
    w[lo.OFF_INV_STD:lo.OFF_GATE] = 1.0 / X.std(0)
    w[:lo.N_PARAMS] += rng.normal(0, 0.05, lo.N_PARAMS).astype(np.float32)
    g_dev, mse_dev = learn.gradient(records, GAMMA, w, lib=lib)
    g_ref = lo.gradient(w.astype(np.float64), X, Y)
    mse_ref = lo.loss(w.astype(np.float64), X, Y)
    # FP32 forward/backward summed over the records: relative to the
    # largest gradient component
    err = np.max(np.abs(g_dev - g_ref)) / np.max(np.abs(g_ref))
    assert err < 2e-4, err
    assert abs(mse_dev - mse_ref) <= 1e-4 * mse_ref
    y_dev, mse_pred = learn.predict(records, GAMMA, w, lib=lib)
    y_ref, _ = lo.forward(w.astype(np.float64), X)
    assert np.max(np.abs(y_dev[ok] - y_ref)) < 1e-5 * max(1.0, np.max(np.abs(y_ref))), \
        np.max(np.abs(y_dev[ok] - y_ref))
    assert abs(mse_pred - mse_ref) <= 1e-4 * mse_ref
    return err


def reference_start_mse(records):
    """Mean squared error of the reference (acoustic) start in the network's
    target, i.e. of the correction 0, over the learnable records."""
    ok = lo.learnable(records, GAMMA)
    return float(np.mean(lo.features(records[ok], GAMMA)[1] ** 2))


def check_newton_stats(lib, records, w):
    out, summary = learn.newton_stats(records, GAMMA, PMIN, w, lib=lib)
    # the reference start reproduces the sweep's own solve
    assert np.array_equal(out[:, 3], records[:, 6])
    assert np.array_equal(out[:, 0], records[:, 8])
    assert np.array_equal(out[:, 6], records[:, 7])
    # from the learned start Newton meets the same tolerance (1e-14 on the
    # last update); a restart reproduces the reference solve exactly; where
    # the start is the reference one (gate, floor) so is everything
    rel = np.abs(out[:, 4] - out[:, 3]) / out[:, 3]
    assert rel.max() < 1e-12, rel.max()
    rs = out[:, 2] == 1
    assert np.array_equal(out[rs, 4], out[rs, 3])
    same = out[:, 5] == out[:, 6]
    assert np.array_equal(out[same, 4], out[same, 3]) and np.array_equal(out[same, 1], out[same, 0])
    return summary


def learned_vs_oracle(lib, g, glob, w, steps, cfl=0.4):
    """`steps` device-resident steps with the learned start, and the same
    steps on the oracle (each with its own Courant dt, both periodic).
    Returns (oracle grid, learned grid)."""
    ref = g.copy()
    oracle_steps("port", ref, glob, steps, cfl)
    h = g.copy()
    dg = device.DeviceGrid(h, glob, lib=lib)
    try:
        dg.set_learned(w)
        for cycle in range(steps):
            dg.periodic_boundary()
            dg.pressure_and_dt(cfl)
            dg.step(cycle)
        assert dg.download() == 1
    finally:
        dg.close()
    return ref, h


def check_sweep_uses_network(lib, n=12, seed=13):
    """The sweep's learned start takes the network's output: a network whose
    output is the constant c = 0.7 (every weight zero but the output bias)
    moves Newton's start at every interface that passes the gate, so with
    gate 0 one step differs from the reference start's in the last bits
    (and agrees to the one-step tolerance), and with a gate no interface
    passes it is bitwise the reference start's step."""
    g0, glob, _ = problems.turbulent(n, seed=seed)

    def run(w):
        h = g0.copy()
        dg = device.DeviceGrid(h, glob, lib=lib)
        try:
            if w is not None:
                dg.set_learned(w)
            dg.periodic_boundary()
            dg.pressure_and_dt(0.4)
            dg.step(0)
            assert dg.download() == 1
        finally:
            dg.close()
        return h

    w = np.zeros(learn.N_FLOATS, np.float32)
    w[learn.OFF_B3] = 0.7
    w[learn.OFF_INV_STD:learn.OFF_GATE] = 1.0
    base = run(None)
    w[learn.OFF_GATE] = 0.0
    moved = run(w)
    w[learn.OFF_GATE] = 1.0e30
    closed = run(w)
    for a, b in zip(closed.fields, base.fields):
        assert np.array_equal(a, b)
    w[learn.OFF_GATE] = np.inf    # the closed gate: the plain kernels
    for a, b in zip(run(w).fields, base.fields):
        assert np.array_equal(a, b)
    assert any(not np.array_equal(a, b) for a, b in zip(moved.fields, base.fields))
    check_learned_run(base, moved, one_step=True)


def best_open_gate(w, table):
    """w with the open (finite) gate of fit_gate's table that gave the
    fewest mean updates."""
    w = np.array(w, dtype=np.float32, copy=True)
    w[learn.OFF_GATE] = min((c for c in table if np.isfinite(c)), key=lambda c: table[c][0])
    return w


def check_gate_choice(fitted, table, rel_tol=0.01):
    """fit_gate's rule: the largest gate within rel_tol of the fewest mean
    updates, the closed gate (+inf, the reference start) included."""
    best = min(v[0] for v in table.values())
    ok = [c for c, v in table.items() if v[0] <= best * (1.0 + rel_tol)]
    assert fitted[learn.OFF_GATE] == np.float32(max(ok))  # the weights hold float32
    assert table[float("inf")][1] == 0.0


def check_sweep_network_is_kernels(lib, g, glob, w):
    """The sweep's learned start (the workgroup's evaluation of the network
    on the gated interfaces, from the device copy of the weights) is the
    per-interface network of the statistics kernel (mlp_eval), bit for bit:
    harvesting with the learned start on, every record's start, Newton
    updates, restart flag and p* equal what hipPPMLearnNewtonStats computes
    from the record's inputs.  Returns the fraction of records whose start
    the network moved."""
    w = np.array(w, dtype=np.float32, copy=True)
    w[learn.OFF_GATE] = 0.0
    rec, seen = learn.harvest(g, glob, steps=1, lib=lib, weights=w)
    assert seen == len(rec) > 0
    out, _ = learn.newton_stats(rec, GAMMA, PMIN, w, lib=lib)
    assert np.array_equal(out[:, 5], rec[:, 7])
    assert np.array_equal(out[:, 1], rec[:, 8])
    assert np.array_equal(out[:, 2], rec[:, 9])
    assert np.array_equal(out[:, 4], rec[:, 6])
    moved = float(np.mean(out[:, 5] != out[:, 6]))
    assert moved > 0.5, moved
    return moved


def check_level_uses_network(lib, grids, glob):
    """check_sweep_uses_network for the batched level (hipPPMLevel*): the
    constant network moves the step with gate 0 and leaves it bitwise the
    reference start's with a closed gate."""
    from enzo_ppm.level import Level

    def run(w):
        lv = Level([g.copy() for g in grids], glob, lib=lib)
        try:
            if w is not None:
                lv.set_learned(w)
            lv.step(0, 0.0, 0.4)
            assert lv.status()[0] == 1
            lv.download()
            return lv.grids
        finally:
            lv.close()

    w = np.zeros(learn.N_FLOATS, np.float32)
    w[learn.OFF_B3] = 0.7
    w[learn.OFF_INV_STD:learn.OFF_GATE] = 1.0
    base = run(None)
    w[learn.OFF_GATE] = 0.0
    moved = run(w)
    w[learn.OFF_GATE] = 1.0e30
    closed = run(w)
    w[learn.OFF_GATE] = np.inf
    for gb, gi in zip(base, run(w)):
        for a, b in zip(gb.fields, gi.fields):
            assert np.array_equal(a, b)
    differs = False
    for gb, gm, gc in zip(base, moved, closed):
        for a, b, c in zip(gb.fields, gm.fields, gc.fields):
            assert np.array_equal(c, a)
            differs |= not np.array_equal(b, a)
            assert l1_rel(b, a) <= 1e-11
    assert differs


def l1_rel(a, b):
    """L1 relative difference; absolute where the reference field is zero."""
    den = float(np.sum(np.abs(b)))
    num = float(np.sum(np.abs(a - b)))
    return num / den if den > 0 else num


def check_learned_run(ref, learned, one_step):
    """One step: max relative difference <= 1e-11 in every field; any
    number of steps: L1 relative difference <= 1e-9 per field, and the
    same total mass to round-off.  Returns the per-field L1 differences."""
    rep = field_report(learned, ref)
    if one_step:
        assert all(r[2] <= 1e-11 for r in rep), rep
    l1 = [l1_rel(x, y) for x, y in zip(learned.fields, ref.fields)]
    assert max(l1) <= 1e-9, l1
    sl = learned.active_slice()
    m0, m1 = ref.fields[0][sl].sum(), learned.fields[0][sl].sum()
    assert abs(m1 - m0) <= 1e-13 * abs(m0)
    return l1


def turbulent(n, seed):
    return problems.turbulent(n, seed=seed)
