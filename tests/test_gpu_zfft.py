"""Register-resident z transforms (nz = 256 ... 2048): fourxz against the CPU
FP64 oracle's c2r / r2c, whole RK3 steps against the oracle, and the same
steps against the LDS Stockham kernels (OTBL_Z_STOCKHAM)."""
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


@pytest.fixture(scope="module", params=[256, 512, 1024, 2048])
def zbox(request, oracle_lib):
    nz = request.param
    cfg = config.small(name=f"z{nz}", nx=16, ny=40, nz=nz, kzmax=min(40, nz // 3 - 1))
    st = ic.make_state(cfg)
    return cfg, st


def test_fourxz_long(zbox, oracle_lib):
    cfg, st = zbox
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    assert gb.paths()["z"] == f"engine{cfg.nz}", gb.paths()
    rng = np.random.default_rng(cfg.nz)
    nrows = 13  # not a multiple of the rows per workgroup
    fou = rng.standard_normal((nrows, cfg.nz1))
    L = oracle_lib.lib()
    phys_o = np.zeros((nrows, cfg.nz))
    for r in range(nrows):
        L.oc_c2r(ob.h, optr(np.ascontiguousarray(fou[r])), optr(phys_o[r]))
    dfou, dphys = DevArray.from_host(fou), DevArray(nrows * cfg.nz)
    gb.op("fourxz", dfou.ptr, dphys.ptr, 1, nrows)
    e1 = rel(dphys.to_host().reshape(nrows, cfg.nz), phys_o)
    phys = rng.standard_normal((nrows, cfg.nz))
    fou_o = np.zeros((nrows, cfg.nz1))
    for r in range(nrows):
        L.oc_r2c(ob.h, optr(np.ascontiguousarray(phys[r])), optr(fou_o[r]))
    dphys2, dfou2 = DevArray.from_host(phys), DevArray(nrows * cfg.nz1)
    gb.op("fourxz", dfou2.ptr, dphys2.ptr, -1, nrows)
    e2 = rel(dfou2.to_host().reshape(nrows, cfg.nz1), fou_o)
    gb.close()
    assert e1 < OP_TOL and e2 < OP_TOL, (e1, e2)


def test_step_long_z(zbox, oracle_lib):
    cfg, st = zbox
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.step()
    gb.step(1)
    errs = state_errors(gb, ob)
    eo, eg = ob.ener(), gb.ener()
    gb.close()
    assert_state(errs)
    assert rel(eg, eo) < 1e-11


def test_engine_matches_stockham(zbox):
    cfg, st = zbox
    gf = GpuBox(cfg, st)
    os.environ["OTBL_Z_STOCKHAM"] = "1"
    try:
        gs = GpuBox(cfg, st)
    finally:
        del os.environ["OTBL_Z_STOCKHAM"]
    assert gf.paths()["z"].startswith("engine") and gs.paths()["z"].startswith("stockham")
    gf.step(2)
    gs.step(2)
    errs = {n: rel(gf.get(n), gs.get(n)) for n in ("u", "v", "w", "p")}
    gf.close()
    gs.close()
    assert all(e < STEP_TOL for e in errs.values()), errs


@pytest.fixture(scope="module", params=[257, 1025, 1536, 1537, 2049], ids=["M512", "M2048", "M3072", "M3072max",
                                                                           "M4096"])
def xbox(request, oracle_lib):
    nx = request.param
    cfg = config.small(name=f"x{nx}", nx=nx, ny=40, nz=24, Lx=8.0 * nx / 32)
    st = ic.make_state(cfg)
    return cfg, st


@pytest.mark.parametrize("iopt", [-1, 1])
def test_cosftx_long(xbox, oracle_lib, iopt):
    """Poisson x transforms on the engine's Bluestein lengths (powers of two
    and 3 x 1024) against the oracle's DCT (itself pinned against scipy)."""
    cfg, st = xbox
    gb = GpuBox(cfg, st)
    want = next(m for m in (512, 1024, 2048, 3072, 4096) if m >= 2 * (cfg.nx - 1) - 1)
    assert gb.paths()["dct"] == f"engine{want}", gb.paths()
    f = np.ascontiguousarray(np.random.default_rng(cfg.nx + iopt).standard_normal((cfg.nx, cfg.ny, cfg.nz1)))
    lines = np.ascontiguousarray(f[1:, :cfg.ny - 1, :].transpose(1, 2, 0))
    oracle_lib.lib().oc_cosft(cfg.nx - 1, lines.shape[0] * lines.shape[1], optr(lines), iopt)
    din, dout = DevArray.from_host(f), DevArray(f.size)
    gb.op("cosftx", din.ptr, dout.ptr, iopt)
    out = dout.to_host().reshape(f.shape)[1:, :cfg.ny - 1, :].transpose(1, 2, 0)
    gb.close()
    assert rel(out, lines) < OP_TOL


def test_step_long_x(xbox, oracle_lib):
    """A whole FP64 step of boxes whose Poisson x transforms take each of the
    engine's Bluestein lengths, against the oracle."""
    cfg, st = xbox
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.step()
    gb.step(1)
    errs = state_errors(gb, ob)
    gb.close()
    assert_state(errs, nx=cfg.nx)
