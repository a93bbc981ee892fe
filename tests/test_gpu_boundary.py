"""Domain-face ghost fill on the GPU (hipPPMExternalBoundary) against the
oracle's restatement of ExternalBoundary::SetExternalBoundary, every field
bitwise: reflecting, outflow, inflow, shearing and periodic faces, a grid
inside the domain (only its faces on the domain's edge are filled), one
active cell between reflecting faces, and a 2-D grid; and the same for every
grid of a level of siblings (hipPPMLevelExternalBoundary)."""
import pytest

import boundary_checks as bc
from enzo_ppm import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    L = abi.load_library()
    if L.hipPPMDeviceCount() < 1:
        pytest.fail("HIP library loaded but no GPU is visible")
    return L


@pytest.mark.parametrize("name", sorted(bc.EXT_CASES))
def test_external_boundary_matches_oracle(lib, name):
    assert bc.run_external_case(lib, name) > 0


@pytest.mark.parametrize("types", [[0, 1, 3, 3, 2, 0], [1, 1, 0, 0, 4, 4]])
def test_level_external_boundary_matches_oracle(lib, types):
    bc.run_level_external(lib, types)
