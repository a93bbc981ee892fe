"""AMR flux correction on the GPU (hipPPMCorrectForRefinedFluxes) against the
oracle's restatement of grid::CorrectForRefinedFluxes, every field bitwise,
plus the grid's own boundary fluxes: an interior subgrid, one whose
corrections empty cells (undone), and subgrids on the grid's own faces."""
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
