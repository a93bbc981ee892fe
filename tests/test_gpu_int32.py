"""The integers-32 library (libenzo_ppm_hip_i32.so: enzo_int = int32, the
reference's CONFIG_INTEGERS = 32 build, Make.settings.test-suite:2) on the
GPU: a one-call SolvePPM_DE with subgrid fluxes, device-resident steps with
gravity and per-cell widths, a batched level and the external-boundary
fill, run in a child process bound to the 32-bit ABI, are bitwise the
results of the default 64-bit library, whose results match the oracle."""
import os

import numpy as np
import pytest

import oracle
from enzo_ppm import abi, problems
from enzo_ppm.grid import SubgridFluxes
from helpers import HERE, ROOT
from test_int32_abi import LIB32, child

pytestmark = pytest.mark.gpu


def test_int32_library_matches_int64(tmp_path):
    import sys
    sys.path.insert(0, HERE)
    import int32_cases
    lib = abi.load_library()
    ref = int32_cases.run_cases(lib)
    out = str(tmp_path / "i32.npz")
    print("test_gpu_int32: child process on", LIB32, flush=True)
    child([os.path.join(HERE, "int32_cases.py"), LIB32, out], timeout=300)
    got = dict(np.load(out))
    assert int(got.pop("integer_bits")[0]) == 32 and int(ref.pop("integer_bits")[0]) == 64
    assert bytes(got.pop("version")).decode().endswith("integers-32")
    assert set(got) == set(ref)
    for k in ref:
        assert np.array_equal(np.asarray(ref[k]).view(np.uint8), np.asarray(got[k]).view(np.uint8)), k
    assert ref["solve_status"][0] == 1 and ref["steps_status"][0] == 1
    assert ref["level_status"][0] == 1 and ref["boundary_status"][0] == 1
    # the one-call case against the oracle (the other cases have their own
    # oracle tests through the 64-bit library)
    g, glob, _ = problems.random_smooth([14, 12, 10], seed=31, ncolours=2, flags=(1, 1, 1))
    g.fill_periodic()
    g.dt = 2.0e-3
    fl = SubgridFluxes(g, [((5, 5, 5), (8, 9, 7)), ((3, 3, 3), (16, 14, 12))])
    st, nfb = oracle.solve_ppm_de("port", g, glob, cycle=1, fluxes=fl)
    assert (st, nfb) == tuple(int(x) for x in got["solve_status"])
    for k, f in enumerate(g.fields):
        assert np.array_equal(f.view(np.int64), got[f"solve_field{k}"].view(np.int64)), k
    assert np.array_equal(fl.flat().view(np.int64), got["solve_fluxes"].view(np.int64))
