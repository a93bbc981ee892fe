"""The learned two-shock start on the GPU: harvest of the plain solver's
interface states (p* bit-equal to the reference twoshock.F), the FP32
training kernels' gradients against the float64 oracle, training on the
GPU, the Newton statistics of the trained start on states it was not
trained on, and learned-start steps against the oracle (one step: max
relative difference <= 1e-11; 20 steps: L1 <= 1e-9 per field)."""
import json

import numpy as np
import pytest

import learn_checks as lc
from enzo_ppm import abi, learn, problems
from enzo_ppm.level import Level
from helpers import level_grids, level_reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


@pytest.fixture(scope="module")
def harvested(lib):
    """Plain-solver states of a 64^3 Sod tube (one step) and a 64^3 Mach-3
    turbulent box (two steps)."""
    g, glob, _ = problems.sod_x(64)
    r1, s1 = lc.harvest(lib, g, glob, steps=1)
    assert s1 == lc.swept_interfaces(g) == len(r1)
    g, glob, _ = problems.turbulent(64, seed=5)
    r2, s2 = lc.harvest(lib, g, glob, steps=2, keep=2)
    assert s2 == len(r2)
    return np.concatenate([r1, r2])


@pytest.fixture(scope="module")
def trained(lib, harvested):
    """(weights with the network evaluated wherever the acoustic start is
    above the floor -- gate 0, so every learned-start test exercises it --,
    the training loss per epoch, and the fitted gate's weights)."""
    w, mse = learn.train(harvested, lc.GAMMA, epochs=8, batch=4096, lr=3e-3, seed=2, lib=lib)
    fitted, table = learn.fit_gate(harvested[::8], lc.GAMMA, lc.PMIN, w, lib=lib)
    print("gate", fitted[learn.OFF_GATE], "(mean updates, evaluated fraction) by gate", table)
    lc.check_gate_choice(fitted, table)
    w = w.copy()
    w[learn.OFF_GATE] = 0.0
    return w, mse, fitted, table


@pytest.fixture(scope="module")
def held_out(lib):
    g, glob, _ = problems.turbulent(64, seed=77)
    rec, _ = lc.harvest(lib, g, glob, steps=1, keep=4)
    return rec


def test_harvest_pstar_is_reference_twoshock(harvested):
    lc.check_harvest_pstar(harvested)


def test_gradient_against_oracle(lib, harvested):
    rng = np.random.default_rng(0)
    lc.check_gradient(lib, harvested[rng.choice(len(harvested), 20000, replace=False)])


def test_training_lowers_loss(trained, harvested):
    w, mse = trained[:2]
    ref = lc.reference_start_mse(harvested)
    print("mse per epoch", mse.tolist(), "reference start", ref)
    assert np.all(np.isfinite(w)) and mse[-1] < 0.5 * mse[0] and mse[-1] < ref


def test_newton_gain_on_held_out_states(lib, trained, held_out):
    """What the trained network gains at its best open gate: Newton updates
    per interface from the learned start against the reference start, on
    states of a field it was not trained on (no more updates than the
    reference start), with p* unchanged to the iteration's tolerance."""
    s = lc.check_newton_stats(lib, held_out, lc.best_open_gate(trained[0], trained[3]))
    print(json.dumps(s))
    assert s["newton_updates_learned_start"]["mean"] <= s["newton_updates_reference_start"]["mean"]
    assert s["restarted_fraction"] < 0.01


def test_newton_stats_on_sedov_and_cosmological_states(lib, trained):
    """The same on the AMR level's Sedov blasts and the cosmological Mach-100
    kind of state, neither of them in the training set: p* unchanged to the
    tolerance, few restarts, and at most 2% more mean updates than the
    reference start at the best open gate (measured in the bench's
    `learned.held_out_newton`: within 1.5% either way)."""
    w = lc.best_open_gate(trained[0], trained[3])
    grids, glob, _ = problems.amr_level(16, seed=41, nmin=8, nmax=48, ncolours=0)
    rec = np.concatenate([learn.harvest(g, glob, steps=1, boundary=False, lib=lib)[0]
                          for g in grids])
    g, globc, _ = problems.cosmological(32, seed=42)
    recc, _ = learn.harvest(g, globc, steps=1, lib=lib)
    for name, r in (("sedov", rec), ("cosmological", recc)):
        s = lc.check_newton_stats(lib, r, w)
        print(name, json.dumps(s))
        ref = s["newton_updates_reference_start"]["mean"]
        assert s["newton_updates_learned_start"]["mean"] <= 1.02 * ref, name
        assert s["restarted_fraction"] < 0.03, name


def test_sweep_uses_network(lib):
    lc.check_sweep_uses_network(lib, n=32)


def test_sweep_network_is_kernels(lib, trained):
    g, glob, _ = problems.turbulent(32, seed=5)
    lc.check_sweep_network_is_kernels(lib, g, glob, trained[0])


def test_level_uses_network(lib):
    grids, glob = level_grids("blast", 24)
    lc.check_level_uses_network(lib, grids, glob)


def test_learned_one_step_against_oracle(lib, trained):
    g, glob, _ = problems.turbulent(64, seed=31)
    ref, got = lc.learned_vs_oracle(lib, g, glob, trained[0], steps=1)
    lc.check_learned_run(ref, got, one_step=True)


def test_learned_20_steps_against_oracle(lib, trained):
    g, glob, _ = problems.turbulent(48, seed=32)
    ref, got = lc.learned_vs_oracle(lib, g, glob, trained[0], steps=20)
    l1 = lc.check_learned_run(ref, got, one_step=False)
    print("L1 per field", l1)


def test_learned_sod_against_oracle(lib, trained):
    g, glob, _ = problems.sod_x(64)
    ref, got = lc.learned_vs_oracle(lib, g, glob, trained[0], steps=1)
    lc.check_learned_run(ref, got, one_step=True)


def test_level_learned(lib, trained):
    grids, glob, _ = problems.amr_level(24, seed=3, nmin=8, nmax=64, ncolours=0)
    ref = [g.copy() for g in grids]
    level_reference(ref, glob, 2)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
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


def test_level_learned_blast(lib, trained):
    grids, glob = level_grids("blast", 4)
    ref = [g.copy() for g in grids]
    level_reference(ref, glob, 2)
    lv = Level([g.copy() for g in grids], glob, lib=lib)
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


@pytest.mark.parametrize("steps", [1, 20])
def test_cosmological_level_learned_against_oracle(lib, trained, steps):
    """The cosmological workload's kind of state (comoving at a = 0.1,
    lognormal density sigma 2, Mach 100, static acceleration, 6 species) on
    a 32^3 grid, batched with the learned start evaluated wherever the
    acoustic start is above the pressure floor (gate 0), against the
    oracle's comoving per-grid loop: one step max relative difference
    <= 1e-11 per field, 20 steps L1 <= 1e-9 per field."""
    from enzo_ppm import comoving
    from helpers import oracle_comoving_level
    w = trained[0].copy()
    w[learn.OFF_GATE] = 0.0
    g, glob, _ = problems.cosmological(32, seed=33)
    model = comoving.EdS(0.1, 1.0)
    ref = g.copy()
    oracle_comoving_level([ref], glob, model, 1.0, steps, 0.4)
    lv = Level([g.copy()], glob, lib=lib)
    try:
        lv.set_comoving(True)
        lv.set_learned(w)
        t = 1.0
        for c in range(steps):
            t, _ = comoving.level_step(lv, model, t, c, 0.4)
            assert lv.status()[0] == 1
        lv.download()
    finally:
        lv.close()
    l1 = lc.check_learned_run(ref, lv.grids[0], one_step=steps == 1)
    print(f"cosmological 32^3, {steps} steps: L1 per field", l1)
