"""CPU checks of the C ABI library: it builds, loads, and exports every entry
point include/opentbl.h declares (no compute calls without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "opentbl.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(otbl_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def libpath():
    from opentbl_amd import build, lib
    if not os.path.exists(lib.LIB_PATH):
        build.build()
    return lib.LIB_PATH


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "otbl_create" in syms and "otbl_substep" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (otbl_[a-z0-9_]+)", out))
    missing = [s for s in header_symbols() if s not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_header(libpath):
    from opentbl_amd import lib
    L = lib.lib()
    for s in header_symbols():
        assert s in lib.SIGNATURES, s
        assert hasattr(L, s)


def test_create_rejects_bad_sizes_without_gpu(libpath):
    """Argument validation runs before any device call."""
    import ctypes as C
    import numpy as np
    from opentbl_amd import lib
    L = lib.lib()
    y = np.zeros(100)
    h = C.c_void_p()
    prm = lib.Params(64, 10, 24, 20, 1000.0, 1.0, 1.0, 0.5, 0, 0)   # ny too small
    rc = L.otbl_create(C.byref(prm), y.ctypes.data_as(lib.PD), C.byref(h))
    assert rc == 1 and b"ny" in L.otbl_last_error()
    prm = lib.Params(64, 40, 22, 20, 1000.0, 1.0, 1.0, 0.5, 0, 0)   # nz = 2*11
    y = np.zeros(42)
    rc = L.otbl_create(C.byref(prm), y.ctypes.data_as(lib.PD), C.byref(h))
    assert rc == 1
    prm = lib.Params(64, 40, 24, 64, 1000.0, 1.0, 1.0, 0.5, 0, 0)   # xout = nx
    rc = L.otbl_create(C.byref(prm), y.ctypes.data_as(lib.PD), C.byref(h))
    assert rc == 1 and b"xout" in L.otbl_last_error()


# the flags the reference calls coefb with and their weight counts
# (coeft.f90:589-1023)
COEFB_FLAGS = {0: 3, 1: 5, 3: 4, 4: 4, 5: 4, 6: 4, 7: 4}


def test_coefb_entry_matches_oracle(libpath, oracle_lib):
    """otbl_coefb (host setup arithmetic, no device) against the oracle's
    coefb, which the schemes' polynomial-exactness pins cover, on random
    spacings of both signs (the reference passes negative spacings for the
    closures at the far end of a line)."""
    import ctypes as C
    import numpy as np
    from opentbl_amd import lib
    L, O = lib.lib(), oracle_lib.lib()
    rng = np.random.default_rng(4)
    for flag, n in COEFB_FLAGS.items():
        for _ in range(20):
            sp = rng.uniform(0.2, 2.0, 4) * rng.choice([-1.0, 1.0])
            got, want = np.zeros(5), np.zeros(5)
            assert L.otbl_coefb(n, flag, *sp, got.ctypes.data_as(lib.PD)) == 0, L.otbl_last_error()
            O.oc_coefb(n, flag, *sp, want.ctypes.data_as(C.POINTER(C.c_double)))
            assert np.abs(got[:n] - want[:n]).max() <= 1e-12 * np.abs(want[:n]).max(), (flag, sp)


def test_coefb_entry_rejects_flags_off_the_path(libpath):
    from opentbl_amd import lib
    import numpy as np
    L = lib.lib()
    w = np.zeros(5)
    assert L.otbl_coefb(4, 2, 1.0, 1.0, 1.0, 1.0, w.ctypes.data_as(lib.PD)) == 1
    assert b"flagv" in L.otbl_last_error()
    assert L.otbl_coefb(4, 0, 1.0, 1.0, 1.0, 1.0, w.ctypes.data_as(lib.PD)) == 1   # flag 0 has 3 weights
    assert b"n does not match" in L.otbl_last_error()
