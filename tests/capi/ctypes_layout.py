"""The ctypes mirror's layout in the lines tests/capi/layout.c prints, for the
width this process's enzo_ppm.abi was imported with (ENZO_PPM_INTEGERS).
Test infrastructure only."""
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..",
                                "enzo-dev_amd"))
from enzo_ppm import abi  # noqa: E402

STRUCTS = [("enzo_fluxes", abi.EnzoFluxes), ("enzo_ppm_globals", abi.EnzoPPMGlobals),
           ("enzo_ppm_grid", abi.EnzoPPMGrid), ("hipPPMParameter", abi.HipPPMParameter)]


def lines():
    out = [f"bits {abi.BITS} 0 {C.sizeof(abi.enzo_int)}"]
    for name, cls in STRUCTS:
        out.append(f"{name} - 0 {C.sizeof(cls)}")
        for f, t in cls._fields_:
            out.append(f"{name} {f} {getattr(cls, f).offset} {C.sizeof(t)}")
    return out


if __name__ == "__main__":
    print("\n".join(lines()))
