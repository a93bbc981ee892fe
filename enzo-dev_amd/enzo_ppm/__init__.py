"""enzo_ppm -- MI355X-native PPM "direct Eulerian" hydro step (Enzo's
grid::SolvePPM_DE path) behind a C ABI (include/enzo_ppm.h).

This package is the thin Python side: ctypes bindings of the HIP library
(abi.py), a host grid container mirroring the state SolvePPM_DE reads
(grid.py), seeded synthetic workloads (problems.py) and the device driver
(device.py).  All compute runs in the HIP library; nothing here falls back
to a CPU path.
"""
from . import abi, grid, problems  # noqa: F401
from .grid import Grid, SubgridFluxes, make_globals  # noqa: F401
