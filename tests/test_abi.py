"""The drop-in boundary (CPU only: nothing here launches a kernel)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import ogjk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True,
                         text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads_and_exports_c_abi():
    lib = ogjk.lib()
    assert lib.ogjk_version() >= 100
    syms = _exported(ogjk.LIB_PATH)
    missing = [s for s in ogjk.C_ABI_SYMBOLS if s not in syms]
    assert not missing, missing
    # every function the C header declares is in the list above
    hdr = open(os.path.join(INC, "opengjk", "ogjk.h")).read()
    import re
    declared = set(re.findall(r"\b(ogjk_[a-z0-9_]+)\s*\(", hdr))
    assert declared <= set(ogjk.C_ABI_SYMBOLS), declared - set(ogjk.C_ABI_SYMBOLS)


def test_cpp_dropin_symbols_match_reference_mangling():
    with tempfile.TemporaryDirectory() as d:
        obj = os.path.join(d, "probe.o")
        subprocess.run(["g++", "-std=c++17", "-c", f"-I{INC}", "-o", obj,
                        os.path.join(ROOT, "tests", "abi_probe.cpp")], check=True)
        und = subprocess.run(["nm", "-u", obj], check=True, capture_output=True, text=True).stdout
    need = {line.split()[-1] for line in und.splitlines() if line.strip()}
    assert "_ZN3GJK3GPU16computeDistancesEiPK10gkPolytopeS3_P9gkSimplexPf" in need
    assert "_Z24compute_minimum_distanceiPK10gkPolytopeS1_P9gkSimplexPf" in need
    missing = sorted(need - _exported(ogjk.LIB_PATH))
    assert not missing, missing


@pytest.mark.parametrize("lang,flags", [("c", ["-std=c11"]), ("c", ["-std=c11", "-DOGJK_USE_64BITS"]),
                                        ("c++", ["-std=c++17"]),
                                        ("c++", ["-std=c++17", "-DOGJK_USE_64BITS"])])
def test_layout_static_asserts_compile(lang, flags):
    src = '#include "opengjk/common.h"\nint main(void) { return (int)sizeof(gkSimplex); }\n'
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t." + ("c" if lang == "c" else "cpp"))
        open(f, "w").write(src)
        cc = "gcc" if lang == "c" else "g++"
        subprocess.run([cc, *flags, f"-I{INC}", "-fsyntax-only", f], check=True)


def test_numpy_views_match_abi():
    assert ogjk.SIMPLEX_DTYPE.itemsize == 108 and ogjk.POLYTOPE_DTYPE.itemsize == 32
    import oracle
    assert oracle.simplex_dtype(64).itemsize == 184


def test_invalid_arguments_fail_cleanly():
    """Argument errors are reported before any device work (no GPU needed)."""
    lib = ogjk.lib()
    assert lib.ogjk_collide_host(7, 1, None, None, None, None, None) == 1
    assert b"mode" in lib.ogjk_last_error()
    assert lib.ogjk_collide_host(1, -3, None, None, None, None, None) == 1
    assert lib.ogjk_collide_host(1, 0, None, None, None, None, None) == 0  # n = 0: no-op
    assert lib.ogjk_set_config(24, 32, 0) == 1
    assert lib.ogjk_set_config(16, 8, 0) == 1
    assert lib.ogjk_set_config(16, 32, 0) == 0


def test_broad_phase_and_contact_arguments_fail_cleanly():
    """Argument errors of the broad-phase and contact-record entry points are
    reported before any device work (no GPU needed)."""
    import ctypes as C
    lib = ogjk.lib()
    g = ogjk._Grid(0.0, 10.0, 8, 0)
    m, f = C.c_int32(), C.c_int64()
    assert lib.ogjk_broad_phase(None, 10, C.byref(g), None, 0, C.byref(m), C.byref(f), None, 0,
                                None) == 1
    assert b"cell_size" in lib.ogjk_last_error()
    g = ogjk._Grid(1.0, 10.0, 2000, 0)
    assert lib.ogjk_broad_phase(None, 10, C.byref(g), None, 0, C.byref(m), C.byref(f), None, 0,
                                None) == 4  # 8e9 cells: unsupported
    g = ogjk._Grid(1.0, 10.0, 8, -1)
    assert lib.ogjk_broad_phase(None, 10, C.byref(g), None, 0, C.byref(m), C.byref(f), None, 0,
                                None) == 1
    g = ogjk._Grid(1.0, 10.0, 8, 0)
    assert lib.ogjk_broad_phase(None, 0, C.byref(g), None, 0, C.byref(m), C.byref(f), None, 0,
                                None) == 0 and m.value == 0 and f.value == 0  # no objects
    assert lib.ogjk_expand_pairs(None, 0, None, None, 0, None, 0, C.byref(m), C.byref(f), None,
                                 0, None) == 0
    assert lib.ogjk_expand_pairs(None, -1, None, None, 0, None, 0, C.byref(m), C.byref(f), None,
                                 0, None) == 1
    # contact records: EPA-only mode (needs simplex input) is refused
    assert lib.ogjk_collide_scene_contacts(2, None, None, 5, None, None, None, 0, None) == 1
    assert b"mode" in lib.ogjk_last_error()
    assert lib.ogjk_collide_scene_contacts(3, None, None, 5, None, None, None, 0, None) == 1
    assert lib.ogjk_collide_scene_contacts(3, None, None, 0, None, None, None, 0, None) == 0
