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
from helpers import noalloc_cases, noalloc_solve

pytestmark = pytest.mark.gpu


def test_no_alloc_matches_solve_ppm_de():
    import torch
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
        fields = [torch.from_numpy(f.copy()).to("cuda:0") for f in g.fields]
        accel = [torch.from_numpy(x.copy()).to("cuda:0") for x in g.accel] if g.gravity else None
        torch.cuda.synchronize()
        sb, fb = noalloc_solve(lib, g, glob, cycle, [t.data_ptr() for t in fields],
                               [t.data_ptr() for t in accel] if accel is not None else None)
        torch.cuda.synchronize()
        assert (sb, fb) == (sa, fa), (k, sb, fb, sa, fa, lib.hipPPMLastError())
        assert sa == abi.SUCCESS
        for x, t in zip(a.fields, fields):
            assert np.array_equal(x.view(np.int64), t.cpu().numpy().view(np.int64)), k
