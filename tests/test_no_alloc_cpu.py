"""hipPPMSolvePPM_DE_NoAlloc (the entry point that allocates nothing: its
device buffers are carved from one static array per GPU) in the library's
own source built for the CPU: two calls in a row, the second on a smaller
grid, bitwise equal to hipPPMSolvePPM_DE on the same inputs with the same
status and fallback count; a grid larger than the array is refused."""
import ctypes as C

import numpy as np

from enzo_ppm import abi, device
from helpers import cpu_library, noalloc_cases, noalloc_solve


def test_no_alloc_matches_solve_ppm_de():
    lib = cpu_library()
    for g, glob, cycle in noalloc_cases():
        a, b = g.copy(), g.copy()
        h = C.c_void_p(None)
        try:
            sa = device.solve_ppm_de(a, glob, cycle=cycle, handle=h, lib=lib)
            fa = lib.hipPPMFallbackCount(h)
        finally:
            if h.value:
                lib.hipPPMDestroy(h)
        # on the CPU build "device" memory is host memory: the arrays in place
        sb, fb = noalloc_solve(lib, b, glob, cycle, [f.ctypes.data for f in b.fields],
                               [x.ctypes.data for x in b.accel] if b.accel is not None else None)
        assert (sb, fb) == (sa, fa) and sa == abi.SUCCESS
        for x, y in zip(a.fields, b.fields):
            assert np.array_equal(x.view(np.int64), y.view(np.int64))


def test_no_alloc_refuses_a_grid_beyond_its_array():
    lib = cpu_library()
    from enzo_ppm import problems
    g, glob, _ = problems.random_smooth([4, 4, 4], seed=1, ncolours=3)
    g.dims = [200, 200, 200]   # described as 200^3: 8e6 cells x 19 arrays > 256 MiB
    g.end = [196, 196, 196]
    st, _ = noalloc_solve(lib, g, glob, 0, [f.ctypes.data for f in g.fields])
    assert st == abi.ERR_ARGUMENT
    assert b"scratch" in lib.hipPPMLastError()
