"""hipPPMSolvePPM_DE_NoAlloc on the GPU (the entry point that allocates
nothing: its device buffers are carved from one static array per GPU, found
with hipGetSymbolAddress): called through the C ABI on fields already in
device memory, two calls in a row, the second on a smaller grid, each
bitwise equal to hipPPMSolvePPM_DE's result on the same inputs, with the
same status and fallback count."""
import ctypes as C

import numpy as np
import pytest

from enzo_ppm import abi, device
from helpers import DeviceArrays, noalloc_cases, noalloc_solve

pytestmark = pytest.mark.gpu


def test_no_alloc_matches_solve_ppm_de():
    lib = abi.load_library()
    for k, (g, glob, cycle) in enumerate(noalloc_cases()):
        print(f"no-alloc call {k}: dims {g.dims}, colours {g.ncolours}, gravity {g.gravity}",
              flush=True)
        a = g.copy()
        h = C.c_void_p(None)
        try:
            sa = device.solve_ppm_de(a, glob, cycle=cycle, handle=h, lib=lib)
            fa = lib.hipPPMFallbackCount(h)
        finally:
            if h.value:
                lib.hipPPMDestroy(h)
        # the fields on the device, through the library's own HIP runtime
        fields = DeviceArrays(g.fields)
        accel = DeviceArrays(g.accel) if g.gravity else None
        try:
            sb, fb = noalloc_solve(lib, g, glob, cycle, fields.addrs(),
                                   accel.addrs() if accel is not None else None)
            assert (sb, fb) == (sa, fa), (k, sb, fb, sa, fa, lib.hipPPMLastError())
            assert sa == abi.SUCCESS
            out = np.empty_like(g.fields[0])
            for i, x in enumerate(a.fields):
                fields.read(i, out)
                assert np.array_equal(x.view(np.int64), out.view(np.int64)), (k, i)
        finally:
            fields.free()
            if accel is not None:
                accel.free()
