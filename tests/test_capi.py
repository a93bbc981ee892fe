"""The HIP library builds, loads without a GPU, exports every entry point the
header declares, and fails loudly (no CPU path) when no GPU is visible."""
import ctypes as C
import os
import re

import pytest

from enzo_ppm import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "enzo_ppm.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hipPPM\w+)\s*\(", txt)))


def test_header_declarations_have_prototypes():
    names = declared()
    assert len(names) >= 20
    assert set(names) == set(abi.PROTOTYPES), set(names) ^ set(abi.PROTOTYPES)


def test_library_exports_every_symbol():
    lib = abi.load_library()
    for n in declared():
        assert hasattr(lib, n), n
    assert lib.hipPPMVersion().startswith(b"enzo_ppm")


def test_struct_layouts():
    # struct fluxes: 4 x [3][3] int64 + 2 x [30][3] pointers (Fluxes.h:29-37)
    assert C.sizeof(abi.EnzoFluxes) == 4 * 9 * 8 + 2 * 30 * 3 * 8
    assert C.sizeof(abi.EnzoPPMGlobals) == 16 * 8


def test_no_cpu_fallback_without_gpu():
    lib = abi.load_library()
    if lib.hipPPMDeviceCount() > 0:
        pytest.skip("a GPU is visible")
    from enzo_ppm import device, problems
    g, glob, _ = problems.random_smooth([8, 8, 8], seed=1)
    st = device.solve_ppm_de(g, glob, lib=lib)
    assert st == abi.ERR_NO_DEVICE


def test_bad_arguments_rejected():
    lib = abi.load_library()
    from enzo_ppm import problems
    g, glob, _ = problems.random_smooth([8, 8, 8], seed=1)
    from enzo_ppm.grid import make_globals
    _, gd = make_globals(**glob)
    para = abi.HipPPMParameter()
    # 2 ghost zones only: rejected (the stencil reaches 3 cells out)
    st = lib.hipPPMInitParameter(C.byref(para), 0, 1, 2, 3, 4, 5, 30, abi.larr([0, 0, 0]),
                                 abi.iarr([2, 2, 2]), abi.iarr([9, 9, 9]), abi.iarr([12, 12, 12]),
                                 0.1, 0.1, 0.1, 5 / 3, 0, 1, 1, 0, 0, 0, 1, 1e-3, 0.1, 0, 1e-20, 0, 1,
                                 0, abi.iarr([0]), 3, 1e-30)
    assert st == abi.ERR_ARGUMENT
    # diffusion type 2 (reads undefined values in the reference) not supported
    para2 = g.hip_parameter(lib, gd)
    para2.PPMDiffusionParameter = 2
    h = C.c_void_p(None)
    assert lib.hipPPMInitData(C.byref(h), C.byref(para2), None) in (abi.ERR_NOT_SUPPORTED,
                                                                   abi.ERR_NO_DEVICE)
