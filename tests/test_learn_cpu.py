"""The learned two-shock start through the library's own HIP source built
for the CPU (tests/emu/hipcpu): harvest, the FP32 training kernels against
the float64 oracle, Newton statistics, and learned-start sweeps against the
oracle, at sizes the CPU build finishes in seconds."""
import numpy as np
import pytest

import learn_checks as lc
from enzo_ppm import learn, problems
from enzo_ppm.level import Level
from helpers import cpu_library, level_grids, level_reference


@pytest.fixture(scope="module")
def cpulib():
    return cpu_library()


@pytest.fixture(scope="module")
def harvested(cpulib):
    g, glob, _ = problems.turbulent(12, seed=3)
    rec, seen = lc.harvest(cpulib, g, glob, steps=2)
    return g, glob, rec, seen


@pytest.fixture(scope="module")
def trained(cpulib, harvested):
    _, _, rec, _ = harvested
    w, mse = learn.train(rec, lc.GAMMA, epochs=8, batch=512, lr=1e-2, seed=1, lib=cpulib)
    fitted, table = learn.fit_gate(rec[::3], lc.GAMMA, lc.PMIN, w, lib=cpulib)
    lc.check_gate_choice(fitted, table)
    w = w.copy()
    w[learn.OFF_GATE] = 0.0   # the network wherever the acoustic start is above the floor
    return w, mse, fitted, table


def test_harvest_every_interface_once(harvested):
    g, _, rec, seen = harvested
    assert seen == 2 * lc.swept_interfaces(g) == len(rec)
    assert set(np.unique(rec[:, 9])) == {0.0}


def test_harvest_pstar_is_reference_twoshock(harvested):
    lc.check_harvest_pstar(harvested[2])


def test_harvest_keep_subsamples(cpulib):
    g, glob, _ = problems.turbulent(12, seed=3)
    rec, seen = lc.harvest(cpulib, g, glob, steps=1, keep=8)
    n = lc.swept_interfaces(g)
    assert seen == len(rec) and 0.08 * n < seen < 0.17 * n
    lc.check_harvest_pstar(rec)


def test_harvest_capacity(cpulib):
    g, glob, _ = problems.turbulent(12, seed=3)
    rec, seen = lc.harvest(cpulib, g, glob, steps=1, cap=1000)
    assert len(rec) == 1000 and seen == lc.swept_interfaces(g)
    lc.check_harvest_pstar(rec)


def test_gradient_against_oracle(cpulib, harvested):
    lc.check_gradient(cpulib, harvested[2][:3000])


def test_training_lowers_loss(trained, harvested):
    w, mse = trained[:2]
    assert np.all(np.isfinite(w)) and mse[-1] < 0.5 * mse[0], mse
    # the trained guess beats the reference start in the target's own metric
    assert mse[-1] < lc.reference_start_mse(harvested[2]), (mse[-1],
                                                            lc.reference_start_mse(harvested[2]))


def test_newton_stats(cpulib, trained, harvested):
    # on its own training states, at the best open gate, the network saves
    # Newton updates
    s = lc.check_newton_stats(cpulib, harvested[2], lc.best_open_gate(trained[0], trained[3]))
    assert s["newton_updates_learned_start"]["mean"] < s["newton_updates_reference_start"]["mean"]


def test_learned_steps_against_oracle(cpulib, trained):
    g, glob, _ = problems.turbulent(12, seed=11)
    ref, got = lc.learned_vs_oracle(cpulib, g, glob, trained[0], steps=1)
    lc.check_learned_run(ref, got, one_step=True)
    ref, got = lc.learned_vs_oracle(cpulib, g, glob, trained[0], steps=4)
    lc.check_learned_run(ref, got, one_step=False)


def test_sweep_uses_network(cpulib):
    lc.check_sweep_uses_network(cpulib)


def test_sweep_network_is_kernels(cpulib, trained):
    g, glob, _ = problems.turbulent(12, seed=5)
    lc.check_sweep_network_is_kernels(cpulib, g, glob, trained[0])


def test_level_uses_network(cpulib):
    grids, glob = level_grids("blast", 3)
    lc.check_level_uses_network(cpulib, grids, glob)


def test_untrained_network_still_converges(cpulib):
    """Any weights: Newton from a poor start restarts from the reference
    start where it does not converge, and the step stays within tolerance."""
    g, glob, _ = problems.turbulent(12, seed=12)
    w = learn.init_weights(7)
    w[learn.OFF_B3] = 6.0          # start e^6 times the acoustic one: far off
    ref, got = lc.learned_vs_oracle(cpulib, g, glob, w, steps=1)
    lc.check_learned_run(ref, got, one_step=True)


def test_learned_weights_rejected_when_not_finite(cpulib):
    g, glob, _ = problems.turbulent(12, seed=12)
    from enzo_ppm import device
    dg = device.DeviceGrid(g, glob, lib=cpulib)
    try:
        for k, bad in ((5, np.nan), (5, np.inf), (learn.OFF_GATE, np.nan),
                       (learn.OFF_GATE, -np.inf)):
            w = learn.init_weights(0)
            w[k] = bad
            with pytest.raises(RuntimeError):
                dg.set_learned(w)
        w = learn.init_weights(0)
        w[learn.OFF_GATE] = np.inf    # the closed gate is accepted
        dg.set_learned(w)
    finally:
        dg.close()


def test_level_learned(cpulib, trained):
    grids, glob = level_grids("blast", 3)
    ref = [g.copy() for g in grids]
    level_reference(ref, glob, 2)
    lv = Level([g.copy() for g in grids], glob, lib=cpulib)
    try:
        lv.set_learned(trained[0])
        for s in range(2):
            lv.step(s, 0.0, 0.4)
            assert lv.status()[0] == 1
        lv.download()
    finally:
        lv.close()
    for a, b in zip(lv.grids, ref):
        for x, y in zip(a.fields, b.fields):
            assert lc.l1_rel(x, y) <= 1e-9


def test_learned_work_item_order_and_lds_reads(cpulib, trained, monkeypatch):
    """The learned start's in-tile list (slots taken in work-item order by
    an LDS counter), the workgroup's network evaluation and the acoustic
    quantities kept in dead LDS arrays: with the work-items of every
    barrier interval run forward, in reverse and in two shuffled orders, the
    step is bitwise the same, and no LDS cell is read before the workgroup
    wrote it (every __shared__ array starts as NaN in the CPU build)."""
    import ctypes as C
    from enzo_ppm import device
    cpulib.hipcpu_lds_violations.restype = C.c_ulonglong
    monkeypatch.setenv("HIPCPU_LDS_CHECK", "1")
    monkeypatch.setenv("HIPCPU_LDS_CONTINUE", "1")
    before = cpulib.hipcpu_lds_violations()
    g0, glob, _ = problems.turbulent(12, seed=21)
    out = {}
    for order in ("fwd", "rev", "shuf:1", "shuf:5"):
        monkeypatch.setenv("HIPCPU_ORDER", order)
        g = g0.copy()
        dg = device.DeviceGrid(g, glob, lib=cpulib)
        try:
            dg.set_learned(trained[0])
            for cycle in range(2):
                dg.periodic_boundary()
                dg.pressure_and_dt(0.4)
                dg.step(cycle)
            assert dg.download() == 1
        finally:
            dg.close()
        out[order] = g
    assert cpulib.hipcpu_lds_violations() == before, "LDS read before write"
    for order, g in out.items():
        assert all(np.array_equal(a, b) for a, b in zip(g.fields, out["fwd"].fields)), order
