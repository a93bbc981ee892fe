"""Benchmark of the PPM-DE hydro step on one MI355X.

Metric: FP64 PPM cell-updates per second per GPU (active cells per grid per
step, the reference's definition: Grid_CollectGridInformation.C:36,
EvolveLevel.C:951-957).  A step is one full hydro step on the device-resident
grid: periodic ghost fill, the fused pressure + Courant pass, and the three
cycle-rotated directional sweeps (grid::SolvePPM_DE).  Inputs are resident in
HBM before the timed region; dt stays on the device.

Workload (default): uniform 256^3 grid, 3 ghost zones, FP64, gamma = 5/3,
lognormal density (sigma_lnrho = 0.5), k^-2 Gaussian random velocity field
on k = 1..16 at RMS Mach 3, uniform temperature, three passive colours,
dual energy on, steepening and flattening on, diffusion off, CFL 0.4.

roofline: HBM-bound.  Algorithmic bytes of one sweep launch = active cells x
8 x [(7 + C + G) reads + (6 + C) writes] (C colours, G = 1 with gravity);
achieved = those bytes / the sweep kernel's mean launch time, timed with HIP
events around each launch on the library's stream.
cpu_baseline: the reference Fortran kernels (flang build in oracle/_ref,
driven by the oracle's slice driver on a thread pool over the host CPU share,
at most 16 threads) on one step of the same 256^3 workload, median of 5 runs;
falls back to the oracle's C restatement ("port") when the reference build is
absent.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "enzo-dev_amd"))

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector/matrix peak
# VALU issue: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD every
# 4 cycles (16 lanes per cycle), at the 2.4 GHz peak engine clock
SIMDS, VALU_CYCLES_PER_INST, CLOCK_HZ = 1024, 4, 2.4e9
# rocprofv3 counts of one 256^3 sweep of the turbulent workload with the
# current build (profiles/ppm_sweep_kernel_turbulent256.md): wave-level VALU
# instructions, waves, and the FP64 add / mul / fma / transcendental parts
VALU_PER_SWEEP_256, WAVES_PER_SWEEP_256 = 1.232e9, 345840
F64_ADD, F64_MUL, F64_FMA, F64_TRANS = 0.168e9, 0.210e9, 0.218e9, 0.0355e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="active cells per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: the process's CPU share, at most 16)")
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--level-grids", type=int, default=512,
                    help="grids of the batched AMR-level comparison (0: skip it)")
    ap.add_argument("--level-steps", type=int, default=5)
    ap.add_argument("--learn-epochs", type=int, default=4,
                    help="training passes of the learned start (0: skip that part)")
    ap.add_argument("--learn-records", type=int, default=10_000_000,
                    help="harvested interface states to train on")
    ap.add_argument("--cosmo-n", type=int, default=512,
                    help="active cells per axis of the comoving workload (0: skip it)")
    ap.add_argument("--cosmo-steps", type=int, default=10)
    ap.add_argument("--solver-steps", type=int, default=5,
                    help="timed steps of the same workload with the HLL and HLLC solvers (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of known size (1 GiB read + 1 GiB write, "
                         "3 times) to scale HBM byte counters")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27  # doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": 3}), flush=True)


def host_cpu_share():
    """(threads, nproc, affinity): the CPUs this process may run on
    (sched_getaffinity), at most 16 (the host CPUs that go with one GPU
    of the machine); nproc is os.cpu_count() for the record."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    return max(1, min(aff, 16)), nproc, aff


def cpu_baseline(n, threads, repeats=5):
    """The reference Fortran kernels (oracle/_ref, else the oracle's C
    restatement) on one full step of the n^3 workload (Courant step,
    pressure and the three sweeps), slices of each sweep spread over
    `threads` threads (0: the host CPU share); the median of `repeats` runs,
    each from the same initial state."""
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from enzo_ppm import problems
    share, nproc, aff = host_cpu_share()
    threads = threads or share
    kind = "ref" if oracle.available("ref") else "port"
    g0, glob, run = problems.turbulent(n, seed=2024)
    g0.fill_periodic()
    times = []
    for _ in range(repeats):
        g = g0.copy()
        t0 = time.perf_counter()
        g.dt = oracle.compute_time_step(kind, g, glob, run["cfl"])
        st, _ = oracle.solve_ppm_de(kind, g, glob, cycle=0, nthreads=threads)
        t1 = time.perf_counter()
        if st != 1:
            raise RuntimeError(f"cpu baseline status {st}")
        times.append(t1 - t0)
        del g
    med = statistics.median(times)
    cells = n ** 3
    return dict(value=cells / med, unit="cell-updates/s", cores=threads, nproc=nproc,
                affinity=aff, kind="reference" if kind == "ref" else "port",
                repeats=repeats, seconds=[round(x, 3) for x in times],
                sample=f"one full step (Courant step, pressure, 3 sweeps) of the {n}^3 workload "
                       f"on {threads} threads, median of {repeats} runs: {med:.2f} s")


def level_solver_rates(lib, ngrids, steps, cfl=0.4):
    """The same AMR level batched with RiemannSolver HLL and HLLC (whole
    pencils in one-pencil workgroups sized to the level's pencils): steps
    of the level's Courant dt after two untimed ones."""
    import numpy as np
    from enzo_ppm import problems
    from enzo_ppm.level import Level
    grids, glob, _ = problems.amr_level(ngrids, seed=7, nmin=8, nmax=64, ncolours=0)
    cells = int(sum(int(np.prod(g.active)) for g in grids))
    out = {}
    for name, rs in (("hll", 1), ("hllc", 4)):
        lv = Level([g.copy() for g in grids], dict(glob, RiemannSolver=rs), lib=lib)
        try:
            for c in range(2):
                lv.step(c, 0.0, cfl)
            if lv.status()[0] != 1:
                raise RuntimeError("level step failed")
            t0 = time.perf_counter()
            for c in range(2, 2 + steps):
                lv.step(c, 0.0, cfl)
            st, _, _ = lv.status()
            t1 = time.perf_counter()
            if st != 1:
                raise RuntimeError(f"level step status {st}")
            out[name] = {"value": cells * steps / (t1 - t0), "ms_per_step": (t1 - t0) * 1e3 / steps}
        finally:
            lv.close()
    return out


def level_bench(lib, ngrids, steps, cfl=0.4, weights=None):
    """The AMR-level workload: `ngrids` grids, each axis log-uniform in
    [8, 64] active cells (seeded), smooth sinusoidal density and velocity with
    a Sedov-like blast (pressure x 1e4 inside radius 0.1 of a random centre),
    no colours.  Both forms are device-resident and take the same steps at
    the level's Courant dt: the plain form calls the per-grid entries
    (hipPPMPressureAndTimeStep + hipPPMStep) grid by grid; the batched form
    is one hipPPMLevelStep (graph replay, autotuned tile shapes).  The two
    end states are compared bitwise.  With `weights`, the batched form is
    then timed again with the learned two-shock start (batched+learned)."""
    import numpy as np
    from enzo_ppm import device, problems
    from enzo_ppm.level import Level
    grids, glob, _ = problems.amr_level(ngrids, seed=7, nmin=8, nmax=64, ncolours=0)
    cells = int(sum(int(np.prod(g.active)) for g in grids))
    lv = Level([g.copy() for g in grids], glob, lib=lib)
    dgs = [device.DeviceGrid(g.copy(), glob, lib=lib) for g in grids]
    try:
        lv.step(0, 0.0, cfl)
        st, dt, _ = lv.status()
        if st != 1:
            raise RuntimeError(f"level step status {st}")
        for d in dgs:
            d.pressure_and_dt(cfl, want_host=False)
            d.step(0, dt)
        shapes = lv.tune(1, dt, repeats=2)
        ttimes, (t_plain, t_pers) = lv.tune_times()
        shape_names = ("16x16", "8x32", "4x64")
        classes = ("[0,20]", "[21,40]", "[41,80]", "81+")
        tuner = {"ms_per_step": {classes[c]: {shape_names[s]: round(float(ttimes[c][s]) * 1e3, 3)
                                              for s in range(3) if ttimes[c][s] == ttimes[c][s]}
                                 for c in range(4)},
                 "chosen": {classes[c]: shape_names[shapes[c]] for c in range(4)},
                 "launch_ms_per_step": {"one_workgroup_per_tile": t_plain * 1e3,
                                        "persistent": t_pers * 1e3},
                 "persistent": lv.persistent(),
                 "note": "pencil-length classes (ghosts included), each timed with the other "
                         "classes at their current choice, in class order"}
        lv.step(1, dt, cfl)
        for d in dgs:
            d.pressure_and_dt(cfl, want_host=False)
            d.step(1, dt)
        for d in dgs:
            d.synchronize()
        lv.status()
        t0 = time.perf_counter()
        for c in range(2, 2 + steps):
            for d in dgs:
                d.pressure_and_dt(cfl, want_host=False)
                d.step(c, dt)
        for d in dgs:
            d.synchronize()
        t1 = time.perf_counter()
        for c in range(2, 2 + steps):
            lv.step(c, dt, cfl)
        st, _, _ = lv.status()
        t2 = time.perf_counter()
        if st != 1:
            raise RuntimeError(f"level step status {st}")
        lv.download()
        same = True
        for d, g in zip(dgs, lv.grids):
            if d.download() != 1:
                raise RuntimeError("per-grid step failed")
            for a, b in zip(d.grid.fields, g.fields):
                same &= bool(np.array_equal(a.view(np.int64), b.view(np.int64)))
        batched_learned = None
        if weights is not None:
            lv.set_learned(weights)
            lv.step(2 + steps, dt, cfl)          # graph capture outside the timing
            lv.status()
            t3 = time.perf_counter()
            for c in range(3 + steps, 3 + 2 * steps):
                lv.step(c, dt, cfl)
            st, _, _ = lv.status()
            t4 = time.perf_counter()
            if st != 1:
                raise RuntimeError(f"learned level step status {st}")
            batched_learned = cells * steps / (t4 - t3)
            lv.set_learned(None)
        # what the learned-start path itself costs: a network whose output is
        # zero (every start the acoustic one, so the Newton work is the plain
        # kernels') evaluated wherever the gate passes, against the plain
        # kernels, at several gates
        from enzo_ppm import learn
        wz = np.zeros(learn.N_FLOATS, np.float32)
        wz[learn.OFF_INV_STD:learn.OFF_GATE] = 1.0
        cost = {}
        c = 3 + 2 * steps
        for name, gate in (("plain", None), ("gate_1e30", 1e30), ("gate_1", 1.0),
                           ("gate_0.1", 0.1), ("gate_0", 0.0)):
            if gate is None:
                lv.set_learned(None)
            else:
                wz[learn.OFF_GATE] = gate
                lv.set_learned(wz)
            lv.step(c, dt, cfl)
            c += 1
            lv.status()
            t5 = time.perf_counter()
            for _ in range(steps):
                lv.step(c, dt, cfl)
                c += 1
            if lv.status()[0] != 1:
                raise RuntimeError("learned-path cost step failed")
            cost[name] = (time.perf_counter() - t5) * 1e3 / steps
        lv.set_learned(None)
    finally:
        lv.close()
        for d in dgs:
            d.close()
    plain = cells * steps / (t1 - t0)
    batched = cells * steps / (t2 - t1)
    return {"workload": f"AMR level: {ngrids} grids, axes log-uniform in [8,64] active cells, "
                        "Sedov blast in sinusoidal flow, no colours",
            "cells_per_step": cells, "steps": steps, "dt": dt,
            "plain_per_grid": {"value": plain, "ms_per_step": (t1 - t0) * 1e3 / steps},
            "batched": {"value": batched, "ms_per_step": (t2 - t1) * 1e3 / steps,
                        "tile_shape_by_pencil_length": shapes, "tuner": tuner},
            "unit": "cell-updates/s", "speedup": batched / plain, "bitwise_equal": same,
            **({"batched_learned": {"value": batched_learned,
                                    "ms_per_step": cells / batched_learned * 1e3,
                                    "speedup_vs_plain_per_grid": batched_learned / plain}}
               if batched_learned else {}),
            "learned_path_cost_ms_per_step": {
                **cost, "note": "zero-output network (every start the acoustic one) evaluated "
                                "where the gate passes (1e30: nowhere, 0: everywhere above "
                                "the floor) against the plain kernels"}}


def time_grid(lib, n, steps, warmup, weights=None, solver=None):
    """Device-resident steps of the n^3 turbulent workload: `warmup` untimed,
    then `steps` timed, with HIP events around every sweep launch.  With
    `weights` the sweeps use the learned two-shock start; `solver` sets
    RiemannSolver (1 HLL, 4 HLLC; default two-shock)."""
    from enzo_ppm import device, problems
    g, glob, run = problems.turbulent(n, seed=2024)
    if solver is not None:
        glob = dict(glob, RiemannSolver=solver)
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        if weights is not None:
            dg.set_learned(weights)
        cfl = run["cfl"]
        cycle = 0
        for _ in range(warmup):
            dg.periodic_boundary()
            dg.pressure_and_dt(cfl, want_host=False)
            dg.step(cycle)
            cycle += 1
        dg.synchronize()
        dg.profile(True)
        dg.profile_read()
        dg.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dg.periodic_boundary()
            dg.pressure_and_dt(cfl, want_host=False)
            dg.step(cycle)
            cycle += 1
        dg.synchronize()
        t1 = time.perf_counter()
        sweep_ms, launches = dg.profile_read()
        fb = dg.fallback_count()
    finally:
        dg.close()
    return dict(value=n ** 3 * steps / (t1 - t0), ms_per_step=(t1 - t0) * 1e3 / steps,
                sweep_ms_avg=sweep_ms / max(1, launches), fallback_cells=int(fb), grid=g)


def host_buffer_rate(lib, n, calls=3):
    """The drop-in call with host arrays (hipPPMSolvePPM_DE, as
    grid::SolvePPM_DE would make it): upload of every field, pressure, the
    three sweeps and download, on a handle kept across calls; one untimed
    call, then `calls` timed.  PCIe-inclusive: not the metric."""
    import ctypes as C
    from enzo_ppm import device, problems
    g, glob, run = problems.turbulent(n, seed=2024)
    h = C.c_void_p(None)
    try:
        dg = device.DeviceGrid(g.copy(), glob, lib=lib)
        try:
            dg.periodic_boundary()
            g.dt = dg.pressure_and_dt(run["cfl"])
        finally:
            dg.close()
        elapsed = 0.0
        for c in range(1 + calls):
            g.fill_periodic()                 # the caller's boundary fill, untimed
            t0 = time.perf_counter()
            st = device.solve_ppm_de(g, glob, cycle=c, handle=h, lib=lib)
            t1 = time.perf_counter()
            if st != 1:
                raise RuntimeError(f"SolvePPM_DE status {st}")
            if c > 0:
                elapsed += t1 - t0
    finally:
        if h.value:
            lib.hipPPMDestroy(h)
    return {"value": n ** 3 * calls / elapsed, "unit": "cell-updates/s",
            "ms_per_call": elapsed * 1e3 / calls,
            "what": "hipPPMSolvePPM_DE with host arrays: upload, pressure, 3 sweeps, download"}


def _level_records(lib, ngrids, seed, keep, cap):
    """Plain-solver interface states of an AMR level's grids (Sedov blasts
    in sinusoidal flow, axes 8-64), one step of each grid with its own ghost
    zones (no periodic fill)."""
    import numpy as np
    from enzo_ppm import learn, problems
    grids, glob, _ = problems.amr_level(ngrids, seed=seed, nmin=8, nmax=64, ncolours=0)
    out = []
    for g in grids:
        r, _ = learn.harvest(g, glob, steps=1, keep=keep, cap=cap, boundary=False, lib=lib)
        out.append(r)
    return np.concatenate(out)


def learned_bench(lib, n, steps, warmup, epochs, cap):
    """The learned two-shock start end to end: harvest of plain-solver
    interface states -- the 64^3 Sod tube (one step, every interface), the
    n^3 turbulent box (one step, one interface in `keep` so that about `cap`
    records are kept), the AMR level's Sedov grids (64 of them) and a 128^3
    cosmological box (Mach 100, sigma_lnrho 2; one in 4) --, training on the
    GPU, the gate fitted on a sample of the training states; Newton
    statistics on held-out states of each kind (other seeds); and the n^3
    step timed with the learned start."""
    import numpy as np
    from enzo_ppm import learn, problems
    t0 = time.perf_counter()
    g1, glob1, _ = problems.sod_x(64)
    r1, _ = learn.harvest(g1, glob1, steps=1, lib=lib)
    r3 = _level_records(lib, 64, 11, 1, cap)
    g4, glob4, _ = problems.cosmological(128, seed=12)
    r4, _ = learn.harvest(g4, glob4, steps=1, keep=4, cap=cap, lib=lib)
    del g4
    g2, glob2, _ = problems.turbulent(n, seed=2024)
    swept = 3 * (n + 1) * (n + 6) ** 2
    keep = max(1, -(-swept // max(1, cap - len(r1) - len(r3) - len(r4))))
    r2, _ = learn.harvest(g2, glob2, steps=1, keep=keep, cap=cap, lib=lib)
    del g2
    rec = np.concatenate([r1, r2, r3, r4])
    sources = {"sod64": len(r1), f"turbulent{n}": len(r2), "amr_level": len(r3),
               "cosmological128": len(r4)}
    del r1, r2, r3, r4
    t1 = time.perf_counter()
    gamma = glob2["Gamma"]
    w, mse = learn.train(rec, gamma, epochs=epochs, batch=8192, lr=3e-3, seed=0, lib=lib)
    t2 = time.perf_counter()
    w, gate_table = learn.fit_gate(rec[::16], gamma, 1e-20, w, lib=lib)
    n_train = len(rec)
    del rec
    held = {}
    gh, globh, _ = problems.turbulent(n, seed=4242)
    rh, _ = learn.harvest(gh, globh, steps=1, keep=16, cap=cap, lib=lib)
    del gh
    held[f"turbulent{n}"] = rh
    held["amr_level"] = _level_records(lib, 64, 13, 1, cap)
    gc, globc, _ = problems.cosmological(128, seed=14)
    held["cosmological128"], _ = learn.harvest(gc, globc, steps=1, keep=4, cap=cap, lib=lib)
    del gc
    stats = {}
    w0 = w.copy()
    w0[learn.OFF_GATE] = 0.0
    for k, r in held.items():
        _, s = learn.newton_stats(r, gamma, 1e-20, w, lib=lib)
        _, s0 = learn.newton_stats(r, gamma, 1e-20, w0, lib=lib)
        s["network_everywhere"] = {"newton_updates_learned_start": s0["newton_updates_learned_start"],
                                   "restarted_fraction": s0["restarted_fraction"],
                                   "network_evaluated_fraction": s0["network_evaluated_fraction"]}
        stats[k] = s
    del held
    print("bench: learned-start timing", file=sys.stderr, flush=True)
    t = time_grid(lib, n, steps, warmup, weights=w)
    t.pop("grid")
    return {"harvest": {"records": n_train, "by_source": sources,
                        "sources": f"64^3 Sod (1 step, all interfaces) + {n}^3 turbulent (1 step, "
                                   f"1 in {keep}) + 64 AMR-level Sedov grids (1 step each) + 128^3 "
                                   "cosmological Mach 100 (1 step, 1 in 4)",
                        "seconds": t1 - t0},
            "training": {"epochs": epochs, "batch": 8192, "lr": 3e-3, "seconds": t2 - t1,
                         "mse_per_epoch": [float(x) for x in mse],
                         # +inf: the network cannot save 1% of the updates on the
                         # gate-fit sample, so it is switched off (plain kernels)
                         "gate": (float(w[learn.OFF_GATE]) if np.isfinite(w[learn.OFF_GATE])
                                  else None),
                         "network_on": bool(np.isfinite(w[learn.OFF_GATE])),
                         "gate_choice": {str(k): {"mean_updates": v[0], "evaluated_fraction": v[1]}
                                         for k, v in gate_table.items()}},
            "held_out_newton": stats, "grid_step_learned": t}, w


def cosmo_bench(lib, n, steps, weights, cfl=0.4):
    """The cosmological workload: one n^3 comoving grid (lognormal density
    with sigma_lnrho 2, Mach-100 k^-2 velocities, p = rho T with T = 1 code
    unit, static seeded acceleration, 6 species; Einstein-de Sitter
    expansion from a = 0.1), `steps` comoving steps each (periodic fill,
    Courant with a(t), a(t+dt/2)-scaled sweeps, expansion terms): first the
    batched form (the grid as a level: graph replay, tile shape autotuned),
    without and (given `weights`) with the learned start, then the plain
    per-grid form.  Each form starts from the same state and takes two
    untimed steps first."""
    from enzo_ppm import comoving, device, problems
    from enzo_ppm.level import Level
    say = lambda m: print(f"bench: cosmological {n}^3: {m}", file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    g, glob, _ = problems.cosmological(n, seed=5, log=say)
    gen_s = time.perf_counter() - t0
    model = comoving.EdS(0.1, 1.0)
    cells = n ** 3
    out = {"workload": f"{n}^3 comoving grid, a = 0.1 (EdS), sigma_lnrho 2, Mach 100, "
                       "static acceleration, 6 species, dual energy, periodic",
           "steps": steps, "generation_seconds": gen_s}
    g0 = g.copy()
    batched_state = None
    forms = [("batched", None)] + ([("batched_learned", weights)] if weights is not None else [])
    for name, w in forms:
        say(f"{name} form")
        g = g0.copy()
        lv = Level([g], glob, lib=lib)
        try:
            lv.set_comoving(True)
            t, dt = comoving.level_step(lv, model, 1.0, 0, cfl)
            shapes = lv.tune(1, dt, repeats=1)
            if w is not None:
                lv.set_learned(w)
            t, _ = comoving.level_step(lv, model, t, 1, cfl)   # graph capture
            if lv.status()[0] != 1:
                raise RuntimeError("level step failed")
            t1 = time.perf_counter()
            for c in range(2, 2 + steps):
                t, dt = comoving.level_step(lv, model, t, c, cfl)
            st, _, nfb = lv.status()
            t2 = time.perf_counter()
            if st != 1:
                raise RuntimeError(f"level step status {st}")
            lv.download()
        finally:
            lv.close()
        out[name] = {"value": cells * steps / (t2 - t1), "ms_per_step": (t2 - t1) * 1e3 / steps,
                     "fallback_cells": int(nfb), "a_end": model(t)[0], "tile_shape": shapes[3]}
        if name == "batched":
            batched_state = g
        del lv, g
    say("plain per-grid form")
    g = g0
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        dg.set_comoving(True)
        t, _ = comoving.device_step(dg, model, 1.0, 0, cfl)
        t, _ = comoving.device_step(dg, model, t, 1, cfl)
        dg.synchronize()
        t1 = time.perf_counter()
        for c in range(2, 2 + steps):
            t, dt = comoving.device_step(dg, model, t, c, cfl)
        dg.synchronize()
        t2 = time.perf_counter()
        if dg.download() != 1:
            raise RuntimeError("per-grid comoving step failed")
    finally:
        dg.close()
    out["plain_per_grid"] = {"value": cells * steps / (t2 - t1),
                             "ms_per_step": (t2 - t1) * 1e3 / steps}
    # the batched and per-grid forms took the same steps from the same state
    import numpy as np
    out["bitwise_equal_batched_vs_per_grid"] = all(
        np.array_equal(a.view(np.int64), b.view(np.int64))
        for a, b in zip(batched_state.fields, g.fields))
    del batched_state
    key = "batched_learned" if weights is not None else "batched"
    out["speedup"] = out[key]["value"] / out["plain_per_grid"]["value"]
    # algorithmic bytes per cell-update: 8 x [3 (R + W) + P] with R = 7 + C + G,
    # W = 6 + C, P = 6 (C = 6 species, G = 1)
    B = 8 * (3 * ((7 + 6 + 1) + (6 + 6)) + 6)
    out["roofline"] = {"bytes_per_cell_update": B,
                       "achieved_GBs": out[key]["value"] * B / 1e9,
                       "frac": out[key]["value"] * B / 1e9 / PEAK_HBM_GBS}
    return out


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)
    from enzo_ppm import abi
    lib = abi.load_library()
    if lib.hipPPMDeviceCount() < 1:
        raise RuntimeError("no GPU visible")
    n = a.n
    print(f"bench: {n}^3 workload, plain start", file=sys.stderr, flush=True)
    t = time_grid(lib, n, a.steps, a.warmup)
    g = t["grid"]
    cells = n ** 3
    C = g.ncolours
    G = 1 if g.gravity else 0
    bytes_sweep = cells * 8 * ((7 + C + G) + (6 + C))
    sweep_avg_s = t["sweep_ms_avg"] / 1e3
    achieved = bytes_sweep / sweep_avg_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_turbulent{n}.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("sweep_bytes_per_launch")
    out = {
        "metric": "FP64 PPM cell-updates/s/GPU (batched+learned vs plain), % HBM roofline",
        "value": t["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal density, k^-2 Mach-3 velocity field)",
        "config": {"workload": f"plain per-grid PPM-DE step, uniform {n}^3 turbulent grid, "
                               f"3 colours, dual energy, periodic", "grid": n,
                   "form": "plain per-grid", "fallback_cells": t["fallback_cells"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": "ppm_sweep_kernel", "kernel_ms_avg": sweep_avg_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_sweep},
        "cpu_baseline": None,
    }
    if n == 256:
        # FP64 work of one 256^3 sweep of this workload, counted by rocprofv3
        # (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 x 64 lanes, an FMA two flops;
        # profiles/ppm_sweep_kernel_turbulent256.md) over the live sweep time:
        # the arithmetic side of the kernel's bound
        flops = (F64_ADD + F64_MUL + 2 * F64_FMA + F64_TRANS) * 64
        out["fp64_valu"] = {"flop_per_launch": flops, "achieved_tflops": flops / sweep_avg_s / 1e12,
                            "peak_tflops": PEAK_FP64_TFLOPS,
                            "frac": flops / sweep_avg_s / 1e12 / PEAK_FP64_TFLOPS,
                            "valu_instructions_per_launch": VALU_PER_SWEEP_256,
                            "note": "instruction counts from the committed rocprofv3 pass; "
                                    "time measured here"}
        # the kernel's actual bound: VALU issue.  Interfaces per sweep: 257
        # per pencil (i1 .. i2+1) on every pencil of the 262^2 cross-section
        ifaces = (n + 1) * (n + 6) ** 2
        issue_s = VALU_PER_SWEEP_256 * VALU_CYCLES_PER_INST / SIMDS / CLOCK_HZ
        out["valu_issue"] = {"bound": "valu", "valu_wave_instructions": VALU_PER_SWEEP_256,
                             "waves": WAVES_PER_SWEEP_256,
                             "valu_per_wave": VALU_PER_SWEEP_256 / WAVES_PER_SWEEP_256,
                             "interfaces": ifaces,
                             "valu_wave_instructions_per_interface": VALU_PER_SWEEP_256 / ifaces,
                             "issue_ms": issue_s * 1e3, "kernel_ms": sweep_avg_s * 1e3,
                             "frac": issue_s / sweep_avg_s,
                             "note": "issue time = instructions x 4 cycles / 1024 SIMDs / "
                                     "2.4 GHz; frac = the share of the kernel's time the "
                                     "SIMDs spend issuing VALU instructions"}
    if a.solver_steps > 0:
        # RiemannSolver HLL / HLLC on the same workload: whole pencils in the
        # one-pencil workgroups (not the metric; the two-shock solver is)
        print("bench: HLL and HLLC solvers", file=sys.stderr, flush=True)
        out["other_solvers"] = {}
        for name, rs in (("hll", 1), ("hllc", 4)):
            try:
                r = time_grid(lib, n, a.solver_steps, 1, solver=rs)
                r.pop("grid")
                out["other_solvers"][name] = r
            except Exception as e:  # reported, not fatal
                out["other_solvers"][name] = {"error": str(e)}
    try:
        out["host_buffers"] = host_buffer_rate(lib, n)
    except Exception as e:  # reported, not fatal
        out["host_buffers"] = {"error": str(e)}
    weights = None
    if a.learn_epochs > 0:
        print("bench: learned start (harvest, training, statistics)", file=sys.stderr, flush=True)
        try:
            out["learned"], weights = learned_bench(lib, n, a.steps, a.warmup, a.learn_epochs,
                                                    a.learn_records)
        except Exception as e:  # reported, not fatal
            out["learned"] = {"error": str(e)}
    if a.level_grids > 0:
        print(f"bench: AMR level of {a.level_grids} grids", file=sys.stderr, flush=True)
        try:
            out["level"] = level_bench(lib, a.level_grids, a.level_steps, weights=weights)
        except Exception as e:  # reported, not fatal
            out["level"] = {"error": str(e)}
        if a.solver_steps > 0 and "error" not in out["level"]:
            try:
                out["level"]["other_solvers"] = level_solver_rates(lib, a.level_grids,
                                                                   a.level_steps)
            except Exception as e:  # reported, not fatal
                out["level"]["other_solvers"] = {"error": str(e)}
    if a.cosmo_n > 0:
        try:
            out["cosmological"] = cosmo_bench(lib, a.cosmo_n, a.cosmo_steps, weights)
        except Exception as e:  # reported, not fatal
            out["cosmological"] = {"error": str(e)}
    if not a.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        try:
            out["cpu_baseline"] = cpu_baseline(n, a.cpu_threads, a.cpu_repeats)
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
