"""The integers-32 build of the C ABI (enzo_int = int32, the reference's
CONFIG_INTEGERS = 32, Make.settings.test-suite:2; macros_and_parameters.h:
201-213) on the CPU: struct layouts as a C caller sees them against the
ctypes mirror for both widths, every declared entry point exported with
the width named by hipPPMVersion, the ctypes prototypes against the
header's declarations, and the library's own source built for the CPU with
32-bit integers giving bitwise the results of the 64-bit build."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from enzo_ppm import abi
from helpers import CSRC, HERE, HIP_UNITS, ROOT, cpu_library

HEADER = os.path.join(ROOT, "include", "enzo_ppm.h")
LIB32 = os.path.join(ROOT, "enzo-dev_amd", "enzo_ppm", "libenzo_ppm_hip_i32.so")
CPU_LIB32 = os.path.join(HERE, "emu", "libenzo_ppm_cpu_i32.so")


def child(args, bits=32, **kw):
    env = dict(os.environ, ENZO_PPM_INTEGERS=str(bits))
    return subprocess.run([sys.executable] + args, env=env, check=True, capture_output=True,
                          text=True, **kw)


@pytest.mark.parametrize("bits", [64, 32])
def test_struct_layouts_match_c(bits, tmp_path):
    exe = str(tmp_path / f"layout{bits}")
    flags = ["-DENZO_PPM_INTEGERS_32"] if bits == 32 else []
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror"] + flags + ["-o", exe,
                    os.path.join(HERE, "capi", "layout.c")], check=True)
    c_lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    py_lines = child([os.path.join(HERE, "capi", "ctypes_layout.py")], bits=bits).stdout.split("\n")
    assert c_lines == py_lines
    assert c_lines[0] == f"bits {bits} 0 {bits // 8}"


def declarations():
    """{name: [argument type strings]} of every function the header declares."""
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*([A-Za-z_][\w \*]*?)\b(hipPPM\w+)\s*\(([^)]*)\)\s*;", txt, re.M):
        args = [a.strip() for a in m.group(3).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(2)] = (m.group(1).strip(), args)
    return out


def expected_ctype(decl):
    """The ctypes type a declaration's C type maps to (the binding's rules)."""
    t = re.sub(r"\bconst\b", "", decl).strip()
    arr = t.endswith("]")
    t = re.sub(r"\[\s*\]$", "", t).strip()
    stars = t.count("*") + (1 if arr else 0)
    base = re.sub(r"\*", " ", t).split()[0]
    table = {"enzo_int": abi.I, "enzo_long_int": abi.LI, "enzo_float": C.c_double,
             "int": C.c_int, "hipPPMParameter": abi.HipPPMParameter,
             "enzo_fluxes": abi.EnzoFluxes, "enzo_ppm_grid": abi.EnzoPPMGrid,
             "enzo_ppm_globals": abi.EnzoPPMGlobals}
    if base in ("hipPPMData", "hipPPMLevel", "void") or (base == "float" and stars == 1):
        return C.c_void_p if stars <= 1 else C.POINTER(C.c_void_p)
    if base == "char" and stars == 1:
        return C.c_char_p
    ct = table[base]
    for _ in range(stars):
        ct = C.POINTER(ct)
    return ct


def test_prototypes_follow_the_header():
    decls = declarations()
    assert set(decls) == set(abi.PROTOTYPES)
    for name, (ret, args) in decls.items():
        res, argtypes = abi.PROTOTYPES[name]
        want = [expected_ctype(a) for a in args]
        assert len(argtypes) == len(want), name
        for k, (a, b) in enumerate(zip(argtypes, want)):
            assert a == b, (name, k, args[k], a, b)
        assert res == expected_ctype(ret), (name, ret)


def test_int32_library_exports():
    if not os.path.exists(LIB32):
        pytest.fail(f"{LIB32} is missing: build it first")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from enzo_ppm import abi\n"
            "lib = abi.load_library()\n"
            "print(abi.LIB_NAME, lib.hipPPMVersion().decode())\n"
            % os.path.join(ROOT, "enzo-dev_amd"))
    out = child(["-c", code]).stdout.split()
    assert out[0] == "libenzo_ppm_hip_i32.so"
    assert out[-1] == "integers-32"
    assert abi.load_library().hipPPMVersion().decode().endswith("integers-64")


def cpu_library32():
    """The library's own HIP source built for the CPU with 32-bit integers."""
    src = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".hpp"))]
    src.append(HEADER)
    if not os.path.exists(CPU_LIB32) or os.path.getmtime(CPU_LIB32) < max(map(os.path.getmtime, src)):
        subprocess.run(["g++", "-x", "c++", "-O1", "-std=c++17", "-fPIC", "-shared",
                        "-Wno-attributes", "-U_FORTIFY_SOURCE", "-Wl,-Bsymbolic",
                        "-ffp-contract=off", "-fno-fast-math", "-DENZO_PPM_INTEGERS_32", "-I",
                        os.path.join(HERE, "emu", "hipcpu"), "-o", CPU_LIB32]
                       + [os.path.join(CSRC, f) for f in HIP_UNITS], check=True)
    return CPU_LIB32


def test_int32_cpu_build_matches_int64(tmp_path):
    sys.path.insert(0, HERE)
    import int32_cases
    ref = int32_cases.run_cases(cpu_library())
    out = str(tmp_path / "i32.npz")
    child([os.path.join(HERE, "int32_cases.py"), cpu_library32(), out], timeout=900)
    got = dict(np.load(out))
    assert int(got.pop("integer_bits")[0]) == 32 and int(ref.pop("integer_bits")[0]) == 64
    assert bytes(got.pop("version")).decode().endswith("integers-32")
    assert set(got) == set(ref)
    for k in ref:
        assert np.array_equal(np.asarray(ref[k]).view(np.uint8), np.asarray(got[k]).view(np.uint8)), k
    assert ref["solve_status"][0] == 1 and ref["steps_status"][0] == 1
    assert ref["level_status"][0] == 1 and ref["boundary_status"][0] == 1
