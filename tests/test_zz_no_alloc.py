"""otbl_soltriy_noalloc against otbl_soltriy through the C ABI: the same
solved column and the same factored tridiagonal, bit for bit, on two calls
in a row (the second with fewer rows), with no device allocation of its own
(its scratch is a fixed device array of the library)."""
import ctypes as C

import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import DevArray, GpuBox, check, lib

pytestmark = pytest.mark.gpu


def tridiag(rng, n):
    a = np.zeros((n, 3))
    a[:, 0] = rng.uniform(0.5, 1.0, n)
    a[:, 2] = rng.uniform(0.5, 1.0, n)
    a[:, 1] = -(a[:, 0] + a[:, 2] + 1.0)
    return a


def solve_both(gb, cfg, rng, ny1, ii):
    plane = rng.standard_normal((ny1, cfg.nz1))
    a = tridiag(rng, ny1)
    outs = []
    for name in ("otbl_soltriy", "otbl_soltriy_noalloc"):
        du = DevArray.from_host(np.ascontiguousarray(plane))
        ds = DevArray(plane.size)
        af = np.ascontiguousarray(a.copy())
        check(getattr(lib(), name)(gb.h, du.ptr, ds.ptr, ii, af.ctypes.data_as(C.POINTER(C.c_double)), ny1), name)
        outs.append((ds.to_host().reshape(plane.shape), af))
        du.free()
        ds.free()
    (x0, a0), (x1, a1) = outs
    A = np.diag(a[:, 1]) + np.diag(a[1:, 0], -1) + np.diag(a[:-1, 2], 1)
    x = np.linalg.solve(A, plane[:, 2 * ii] + 1j * plane[:, 2 * ii + 1])
    return x0, a0, x1, a1, x


def test_soltriy_noalloc_matches_soltriy():
    cfg = config.small(name="na", nx=32, ny=40, nz=24)
    gb = GpuBox(cfg, ic.make_state(cfg))
    rng = np.random.default_rng(21)
    for ny1, ii in ((cfg.ny - 1, 2), (13, 0)):
        x0, a0, x1, a1, x = solve_both(gb, cfg, rng, ny1, ii)
        col = x1[:, 2 * ii] + 1j * x1[:, 2 * ii + 1]
        assert np.abs(col - x).max() < 1e-12 * np.abs(x).max()
        assert np.array_equal(x1[:, 2 * ii:2 * ii + 2], x0[:, 2 * ii:2 * ii + 2])
        assert np.array_equal(a1, a0)
    # a size beyond the scratch is refused before any device work
    big = np.zeros((8193, 3))
    d = DevArray(16)
    rc = lib().otbl_soltriy_noalloc(gb.h, d.ptr, d.ptr, 0, big.ctypes.data_as(C.POINTER(C.c_double)), 8193)
    assert rc == 1 and b"8192" in lib().otbl_last_error()
    d.free()
    gb.close()
