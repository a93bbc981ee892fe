"""The HIP sweep's tile code (enzo-dev_amd/csrc/ppm_sweep.hpp), run by the
CPU block emulator with every LDS and global access bounds-checked, against
the oracle: bitwise equality of all fields and captured face fluxes.  This
pins the kernel's indexing, tiling, halo and phase order on the CPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from enzo_ppm import abi, problems
from enzo_ppm.grid import SubgridFluxes, make_globals
from helpers import FALLBACK_CASES, bit_equal, fallback_case, field_report, fill_bc

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu", "libppm_emu.so")


@pytest.fixture(scope="module")
def emu():
    src = os.path.join(HERE, "emu", "ppm_emu.cpp")
    if not os.path.exists(EMU) or os.path.getmtime(EMU) < os.path.getmtime(src):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-fno-fast-math", "-DPPM_EMU_CHECKS", "-o", EMU, src], check=True)
    L = C.CDLL(EMU)
    L.emu_solve_ppm_de.restype = C.c_int
    return L


def emu_solve(L, g, glob, cycle=0, fluxes=None):
    gs, _ = make_globals(**glob)
    cg, keep = g.ctypes_grid()
    cw = g.ctypes_cellwidth()
    colnum = abi.iarr(g.colnum, max(1, len(g.colnum)))
    nsub = fluxes.n if fluxes is not None else 0
    fptr = fluxes.ptrs if fluxes is not None else (C.POINTER(abi.EnzoFluxes) * 1)()
    gstart = fluxes.global_start if fluxes is not None else abi.larr([0, 0, 0])
    cnt = (C.c_int64 * 4)()
    st = L.emu_solve_ppm_de(C.byref(cg), C.byref(gs), C.c_int64(cycle), C.c_int64(nsub), fptr, cw,
                            gstart, C.c_int64(int(g.gravity)), C.c_int64(g.ncolours), colnum,
                            C.c_double(0.0), cnt)
    return st, list(cnt)[:3]


CASES = {
    "sod_diffusion": (lambda: problems.sod_x(16), "periodic", 0.4),
    "random_colours": (lambda: problems.random_smooth([12, 10, 14], seed=1, ncolours=2), "periodic", 0.4),
    "gravity": (lambda: problems.random_smooth([10, 12, 8], seed=2, ncolours=1, gravity=True), "periodic", 0.4),
    "no_dual_energy": (lambda: problems.random_smooth([10, 12, 8], seed=3, dual_energy=False), "periodic", 0.4),
    "diffusion": (lambda: problems.random_smooth([10, 12, 8], seed=4, ncolours=1, flags=(1, 1, 1)), "periodic", 0.4),
    "stretched": (lambda: problems.random_smooth([12, 10, 11], seed=7, ncolours=1, stretch=0.2), "periodic", 0.4),
    "turbulent": (lambda: problems.turbulent(28, seed=3), "periodic", 0.4),
    "sod_1d": (lambda: problems.sod_x(50, rank=1, gamma=1.4, dual_energy=False, flags=(0, 0, 0)), "outflow", 0.9),
    "sod_2d": (lambda: problems.sod_x(20, rank=2, gamma=1.4, flags=(1, 1, 1)), "periodic", 0.9),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_emulated_kernel_matches_oracle(emu, name):
    mk, bc, cfl = CASES[name]
    g, glob, _ = mk()
    a, b = g.copy(), g.copy()
    for cyc in range(3):
        fill_bc(a, bc)
        fill_bc(b, bc)
        a.dt = oracle.compute_time_step("port", a, glob, cfl)
        b.dt = a.dt
        assert oracle.solve_ppm_de("port", a, glob, cycle=cyc)[0] == 1
        st, cnt = emu_solve(emu, b, glob, cycle=cyc)
        assert st == 1 and cnt == [0, 0, 0], (st, cnt)
    rep = field_report(b, a)
    assert all(r[1] for r in rep), rep


def test_emulated_face_fluxes(emu):
    g, glob, _ = problems.random_smooth([20, 18, 16], seed=5, ncolours=2)
    boxes = [((7, 5, 4), (12, 9, 8)), ((3, 3, 3), (22, 20, 18)), ((10, 3, 3), (10, 20, 18))]
    for cyc in range(3):
        g.fill_periodic()
        g.dt = oracle.compute_time_step("port", g, glob, 0.4)
        a, b = g.copy(), g.copy()
        fa, fb = SubgridFluxes(a, boxes), SubgridFluxes(b, boxes)
        assert oracle.solve_ppm_de("port", a, glob, cycle=cyc, fluxes=fa)[0] == 1
        assert emu_solve(emu, b, glob, cycle=cyc, fluxes=fb)[0] == 1
        assert bit_equal(fa.flat(), fb.flat())
        assert all(r[1] for r in field_report(b, a))
        g = a


def test_emulated_table_path(emu, monkeypatch):
    """Coefficient tables instead of the uniform-axis constants."""
    monkeypatch.setenv("PPM_EMU_TABLES", "1")
    g, glob, _ = problems.turbulent(20, seed=6)
    a, b = g.copy(), g.copy()
    a.dt = b.dt = oracle.compute_time_step("port", a, glob, 0.4)
    assert oracle.solve_ppm_de("port", a, glob, cycle=2)[0] == 1
    assert emu_solve(emu, b, glob, cycle=2)[0] == 1
    assert all(r[1] for r in field_report(b, a))


@pytest.mark.parametrize("name", sorted(FALLBACK_CASES))
def test_emulated_fallback(emu, name):
    """Per-cell HLL fallback: the sweep tiles flag the failing pencils and
    one-pencil workgroups of 1024 threads redo them with the reference's
    in-order fallback.  Same status and fallback count as the oracle, and
    the same fields where the step succeeds."""
    g, glob, cycle = fallback_case(name)
    status = FALLBACK_CASES[name][-1]
    a, b = g.copy(), g.copy()
    sa, fa = oracle.solve_ppm_de("port", a, glob, cycle=cycle)
    assert sa == status and fa > 0
    sb, cnt = emu_solve(emu, b, glob, cycle=cycle)
    assert sb == status and cnt[0] == fa, (sb, cnt, fa)
    if status == 1:
        rep = field_report(b, a)
        assert all(r[1] for r in rep), rep
