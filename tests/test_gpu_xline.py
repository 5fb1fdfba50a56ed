"""x-line operators with the windowed backward sweep (lines longer than 80
planes): interpxx and difvisxx against the CPU FP64 oracle on random fields,
including line lengths that leave a partial last block and lines just long
enough for the window, and whole RK3 steps against the oracle and against
the marching kernels (OTBL_X_MARCH)."""
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


@pytest.fixture(scope="module", params=[81, 100, 257])
def xb(request, oracle_lib):
    nx = request.param
    cfg = config.small(name=f"xl{nx}", nx=nx, ny=40, nz=24, Lx=8.0 * nx / 32)
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    assert gb.paths()["x"] == "window", gb.paths()
    yield cfg, st, ob, gb
    gb.close()


def field(cfg, R, seed):
    return np.ascontiguousarray(np.random.default_rng(seed).standard_normal((cfg.nx, R, cfg.nz1)))


@pytest.mark.parametrize("which_u,R_is_u", [(1, True), (0, False), (0, True)])
def test_interpxx_windowed(xb, which_u, R_is_u):
    cfg, _, ob, gb = xb
    R = cfg.ny + 1 if R_is_u else cfg.ny
    g = field(cfg, R, 31)
    out_o = np.zeros_like(g)
    ob.op("interpxx", optr(g), optr(out_o), R, which_u)
    dg, do = DevArray.from_host(g), DevArray(g.size)
    gb.op("interpxx", dg.ptr, do.ptr, R, which_u)
    e = rel(do.to_host().reshape(g.shape), out_o)
    assert e < OP_TOL, e


@pytest.mark.parametrize("which_u,R_is_u", [(1, True), (0, False), (0, True)])
def test_difvisxx_windowed(xb, which_u, R_is_u):
    cfg, _, ob, gb = xb
    R = cfg.ny + 1 if R_is_u else cfg.ny
    prod, vel = field(cfg, R, 32), field(cfg, R, 33)
    out_o = np.zeros_like(prod)
    ob.op("difvisxx", optr(prod), optr(vel), optr(out_o), R, which_u)
    dp, dv, do = DevArray.from_host(prod), DevArray.from_host(vel), DevArray(prod.size)
    gb.op("difvisxx", dp.ptr, dv.ptr, do.ptr, R, which_u)
    e = rel(do.to_host().reshape(prod.shape), out_o)
    assert e < OP_TOL, e


def test_step_windowed(xb, oracle_lib):
    cfg, st, _, _ = xb
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.step()
    gb.step(1)
    errs = state_errors(gb, ob)
    gb.close()
    assert_state(errs)


def test_windowed_matches_marching(xb):
    cfg, st, _, _ = xb
    gw = GpuBox(cfg, st)
    os.environ["OTBL_X_MARCH"] = "1"
    try:
        gm = GpuBox(cfg, st)
    finally:
        del os.environ["OTBL_X_MARCH"]
    assert gw.paths()["x"] == "window" and gm.paths()["x"] == "march"
    gw.step(2)
    gm.step(2)
    errs = {n: rel(gw.get(n), gm.get(n)) for n in ("u", "v", "w", "p")}
    gw.close()
    gm.close()
    assert all(e < STEP_TOL for e in errs.values()), errs


@pytest.mark.parametrize("nx,nseg", [(257, 3), (600, 2), (600, 4)])
def test_segmented_lines(oracle_lib, nx, nseg):
    """Lines split into segments whose forward sweep starts XLINE_LEAD planes
    early from zero: operators and whole steps against one thread per line
    (OTBL_XSEG=1) and the steps against the oracle."""
    cfg = config.small(name=f"xs{nx}", nx=nx, ny=40, nz=24, Lx=8.0 * nx / 32)
    st = ic.make_state(cfg)
    boxes = {}
    for n in (1, nseg):
        os.environ["OTBL_XSEG"] = str(n)
        try:
            boxes[n] = GpuBox(cfg, st)
        finally:
            del os.environ["OTBL_XSEG"]
    g1, gs = boxes[1], boxes[nseg]
    assert g1.paths()["xseg"] == "1" and gs.paths()["xseg"] == str(nseg), gs.paths()
    R = cfg.ny + 1
    g = field(cfg, R, 5)
    q = field(cfg, R, 6)
    for name, args in (("interpxx", (R, 1)), ("difvisxx", (R, 1))):
        outs = []
        for gb in (g1, gs):
            dg, dq, do = DevArray.from_host(g), DevArray.from_host(q), DevArray(g.size)
            if name == "interpxx":
                gb.op(name, dg.ptr, do.ptr, *args)
            else:
                gb.op(name, dg.ptr, dq.ptr, do.ptr, *args)
            outs.append(do.to_host())
        assert rel(outs[1], outs[0]) < 1e-13, name
    ob = oracle_lib.OracleBox(cfg, st)
    ob.step()
    gs.step(1)
    g1.step(1)
    errs = state_errors(gs, ob)
    ediff = {n: rel(gs.get(n), g1.get(n)) for n in ("u", "v", "w")}
    g1.close()
    gs.close()
    assert_state(errs)
    assert all(e < 1e-12 for e in ediff.values()), ediff
