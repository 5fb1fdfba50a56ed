"""Chunked y-line solver on tall columns: the y operators and a full RK3 step
against the CPU FP64 oracle at column lengths that use 20 and 24 rows per
chunk (the lengths of the 1536x315x1024 and 3072x360x2048 boxes), with a
partly filled and an empty last chunk.  The one-thread-per-column kernels
(OTBL_Y_MARCH) agree with it to the same 1e-10 bar (p, the most sensitive
field, differs by a few 1e-12 after two steps from the changed summation
order)."""
import os

import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import DevArray, GpuBox
from parity_util import assert_state, state_errors

pytestmark = pytest.mark.gpu

OP_TOL = 1e-12
STEP_TOL = 1e-10


def rel(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / s)


def optr(a):
    from oracle.oracle import ptr
    return ptr(a)


@pytest.fixture(scope="module", params=[300, 370], ids=["L20", "L24"])
def tall(request, oracle_lib):
    cfg = config.small(name=f"tall{request.param}", nx=16, ny=request.param, nz=24)
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    yield cfg, st, ob, gb
    gb.close()


def field(cfg, R, seed):
    return np.ascontiguousarray(np.random.default_rng(seed).standard_normal((cfg.nx, R, cfg.nz1)))


@pytest.mark.parametrize("loguv", [0, 1])
def test_interpyy_tall(tall, loguv):
    cfg, _, ob, gb = tall
    g = field(cfg, cfg.ny + 1 if loguv == 0 else cfg.ny, 3 + loguv)
    out_o = np.zeros((cfg.nx, cfg.ny + 1, cfg.nz1))
    ob.op("interpyy", optr(g), optr(out_o), loguv, cfg.nx)
    dg, do = DevArray.from_host(g), DevArray(out_o.size)
    gb.op("interpyy", dg.ptr, do.ptr, loguv, cfg.nx)
    out_g = do.to_host().reshape(out_o.shape)
    assert rel(out_g[:, :cfg.ny], out_o[:, :cfg.ny]) < OP_TOL


@pytest.mark.parametrize("which", [0, 1])
def test_differzy_tall(tall, which):
    cfg, _, ob, gb = tall
    g = field(cfg, cfg.ny + 1, 5 + which)
    out_o = np.zeros_like(g)
    ob.op("differzy", optr(g), optr(out_o), which, cfg.nx)
    dg, do = DevArray.from_host(g), DevArray(g.size)
    gb.op("differzy", dg.ptr, do.ptr, which, cfg.nx)
    rows = cfg.ny + 1 if which == 0 else cfg.ny
    out_g = do.to_host().reshape(g.shape)
    assert rel(out_g[:, :rows], out_o[:, :rows]) < OP_TOL


@pytest.mark.parametrize("R_is_u", [True, False])
def test_vistyy_implzy_tall(tall, R_is_u):
    cfg, _, ob, gb = tall
    R = cfg.ny + 1 if R_is_u else cfg.ny
    u = field(cfg, R, 7)
    out_o = np.zeros_like(u)
    ob.op("vistyy", optr(u), optr(out_o), R, cfg.nx)
    du, do = DevArray.from_host(u), DevArray(u.size)
    gb.op("vistyy", du.ptr, do.ptr, R, cfg.nx)
    assert rel(do.to_host().reshape(u.shape), out_o) < OP_TOL
    res = field(cfg, R, 8)
    u_o = u.copy()
    ob.op("implzy", optr(u_o), optr(res), R, -0.021, cfg.nx)
    dr = DevArray.from_host(res)
    gb.op("implzy", du.ptr, dr.ptr, R, -0.021, cfg.nx)
    assert rel(du.to_host().reshape(u.shape), u_o) < OP_TOL


def test_step_tall(tall, oracle_lib):
    cfg, st, _, _ = tall
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.step()
    gb.step(1)
    errs = state_errors(gb, ob)
    gb.close()
    assert_state(errs)


def test_chunked_matches_marching(tall):
    cfg, st, _, _ = tall
    gc = GpuBox(cfg, st)
    os.environ["OTBL_Y_MARCH"] = "1"
    try:
        gm = GpuBox(cfg, st)
    finally:
        del os.environ["OTBL_Y_MARCH"]
    assert gc.paths()["y"].startswith("chunk") and gm.paths()["y"] == "march0"
    gc.step(2)
    gm.step(2)
    for n in ("u", "v", "w", "p"):
        e = rel(gc.get(n), gm.get(n))
        assert e < STEP_TOL, (n, e)
    gc.close()
    gm.close()
