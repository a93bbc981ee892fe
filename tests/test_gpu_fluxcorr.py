"""AMR flux correction on the GPU (hipPPMCorrectForRefinedFluxes) against the
oracle's restatement of grid::CorrectForRefinedFluxes, every field bitwise,
plus the grid's own boundary fluxes: an interior subgrid, one whose
corrections empty cells (undone), and subgrids on the grid's own faces; and
SUBlings (finer grids of a neighbour touching this grid's faces) on either
side, with undone corrections and across a periodic domain boundary."""
import pytest

import fluxcorr_checks as fc
from enzo_ppm import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


@pytest.mark.parametrize("name", sorted(fc.CASES))
def test_flux_correction_matches_oracle(lib, name):
    assert fc.run_case(lib, name) > 0


@pytest.mark.parametrize("r", [2, 4])
def test_flux_projection_matches_oracle(lib, r):
    fc.run_projection(lib, r)


@pytest.mark.parametrize("name", sorted(fc.SUBLING_CASES))
def test_subling_flux_correction_matches_oracle(lib, name):
    assert fc.run_subling_case(lib, name) > 0


@pytest.mark.parametrize("steps", [0, 1])
@pytest.mark.parametrize("name", sorted(fc.CASES))
def test_level_flux_correction_matches_oracle(lib, name, steps):
    print(f"level flux correction {name} steps={steps}", flush=True)
    fc.run_level_case(lib, name, steps)
