"""Cases run through either integer width of the C ABI, for the comparison
of the integers-32 library (enzo_int = int32, the reference's
CONFIG_INTEGERS = 32 build, Make.settings.test-suite:2) with the default
integers-64 one.

The width is fixed per process by ENZO_PPM_INTEGERS (read when enzo_ppm.abi
is imported), so the integers-32 run is a child process:

    ENZO_PPM_INTEGERS=32 python tests/int32_cases.py LIBRARY OUT.npz

`run_cases(lib)` returns {name: array}: the fields, statuses and captured
face fluxes of
  - hipPPMSolvePPM_DE with colours, diffusion and two subgrids' face fluxes,
  - device-resident steps (periodic fill, pressure + Courant, Step) with
    gravity and per-cell widths,
  - a batched level of three grids (hipPPMLevel*, the autotuner included),
  - the external-boundary fill (reflecting / outflow / inflow faces).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "enzo-dev_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def run_cases(lib):
    import ctypes as C

    from enzo_ppm import abi, device, problems
    from enzo_ppm.grid import SubgridFluxes
    from enzo_ppm.level import Level

    out = {"integer_bits": np.array([abi.BITS])}

    # one SolvePPM_DE call: colours, diffusion, flattening, steepening and
    # the face fluxes of two subgrids
    g, glob, _ = problems.random_smooth([14, 12, 10], seed=31, ncolours=2, flags=(1, 1, 1))
    g.fill_periodic()
    g.dt = 2.0e-3
    fl = SubgridFluxes(g, [((5, 5, 5), (8, 9, 7)), ((3, 3, 3), (16, 14, 12))])
    h = C.c_void_p(None)
    try:
        st = device.solve_ppm_de(g, glob, cycle=1, fluxes=fl, handle=h, lib=lib)
        out["solve_status"] = np.array([st, lib.hipPPMFallbackCount(h)])
    finally:
        if h.value:
            lib.hipPPMDestroy(h)
    for k, f in enumerate(g.fields):
        out[f"solve_field{k}"] = f
    out["solve_fluxes"] = fl.flat()

    # device-resident steps with gravity and stretched cell widths
    g, glob, run = problems.random_smooth([12, 10, 14], seed=32, ncolours=1, gravity=True,
                                          stretch=0.1)
    dg = device.DeviceGrid(g, glob, lib=lib)
    dts = []
    try:
        for c in range(3):
            dg.periodic_boundary()
            dts.append(dg.pressure_and_dt(0.4))
            dg.step(c)
        out["steps_status"] = np.array([dg.download()])
    finally:
        dg.close()
    out["steps_dt"] = np.array(dts)
    for k, f in enumerate(g.fields):
        out[f"steps_field{k}"] = f

    # a batched level: three uneven grids, two steps at the level's dt
    grids, glob, _ = problems.amr_level(3, seed=33, nmin=8, nmax=20, ncolours=1)
    lv = Level([x.copy() for x in grids], glob, lib=lib)
    try:
        lv.step(0, 0.0, 0.4)
        # the tuner's choice depends on timings, so only its range is checked
        shapes = lv.tune(1, 1.0e-5, repeats=1)
        assert all(s in (0, 1, 2) for s in shapes), shapes
        lv.step(1, 0.0, 0.4)
        st, dt, nfb = lv.status()
        lv.download()
        out["level_status"] = np.array([st, nfb])
        out["level_dt"] = np.array([dt])
        for i, x in enumerate(lv.grids):
            for k, f in enumerate(x.fields):
                out[f"level_grid{i}_field{k}"] = f
    finally:
        lv.close()

    # external boundary faces: reflecting x, outflow y, inflow z
    g, glob, _ = problems.random_smooth([9, 8, 7], seed=34, ncolours=1)
    dg = device.DeviceGrid(g, glob, lib=lib)
    try:
        inflow = np.arange(g.nfields * 6, dtype=np.float64).reshape(g.nfields, 6) * 0.01 + 0.5
        dg.external_boundary([0, 0, 1, 1, 2, 2], inflow=inflow)
        out["boundary_status"] = np.array([dg.download()])
    finally:
        dg.close()
    for k, f in enumerate(g.fields):
        out[f"boundary_field{k}"] = f
    return out


def main():
    from enzo_ppm import abi
    lib = abi.load_library(sys.argv[1])
    res = run_cases(lib)
    res["version"] = np.frombuffer(lib.hipPPMVersion(), dtype=np.uint8)
    np.savez(sys.argv[2], **res)
    print("int32_cases: done", flush=True)


if __name__ == "__main__":
    main()
