"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Operators are compared on seeded random fields, the full RK3 substep on the
generated boundary-layer state.  Bar: FP64 operators to 1e-12 of the field's
largest magnitude, u, v, w, p after whole steps to 1e-10 relative.
"""
import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import DevArray, GpuBox, OTBL_DT_FROZEN
from parity_util import assert_state, state_errors

pytestmark = pytest.mark.gpu

OP_TOL = 1e-12
STEP_TOL = 1e-10


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    s = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s)


@pytest.fixture(scope="module")
def tiny(oracle_lib):
    cfg = config.small()
    st = ic.make_state(cfg)
    return cfg, st


@pytest.fixture(scope="module")
def boxes(tiny, oracle_lib):
    cfg, st = tiny
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    yield ob, gb
    gb.close()


def rand_field(cfg, R, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((cfg.nx, R, cfg.nz1))
    return np.ascontiguousarray(a)


def optr(a):
    from oracle.oracle import ptr
    return ptr(a)


COEFS = ["inxu", "inxv", "inyu", "inyv", "inbx", "inby", "cofiux", "cofivx", "cofiuy", "cofivy",
         "dcxu", "dcxv", "dcyu", "dcyv", "dcbx", "dcby", "ccon", "cofcxu", "cofcxv", "cofcyu",
         "cofcyv", "vixu", "vixv", "viyu", "viyv", "vyui", "vyvi", "cofvbx", "cofvby", "cofvxu",
         "cofvxv", "cofvyu", "cofvyv", "ayp", "kmod", "kaz", "kaz2", "kvis", "vmagic", "dy",
         "idy", "idyy", "dymin", "hy"]


@pytest.mark.parametrize("name", COEFS)
def test_coefficients(boxes, name):
    ob, gb = boxes
    a = gb.coef(name)
    b = ob.coef(name)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    assert rel(a, b) < 1e-13, (name, rel(a, b))


@pytest.mark.parametrize("which_u,R_is_u", [(1, True), (0, False), (0, True)])
def test_interpxx(boxes, tiny, which_u, R_is_u):
    ob, gb = boxes
    cfg, _ = tiny
    R = cfg.ny + 1 if R_is_u else cfg.ny
    g = rand_field(cfg, R, 11)
    out_o = np.zeros_like(g)
    ob.op("interpxx", optr(g), optr(out_o), R, which_u)
    dg, do = DevArray.from_host(g), DevArray(g.size)
    gb.op("interpxx", dg.ptr, do.ptr, R, which_u)
    assert rel(do.to_host().reshape(g.shape), out_o) < OP_TOL


@pytest.mark.parametrize("which_u,R_is_u", [(1, True), (0, False), (0, True)])
def test_difvisxx(boxes, tiny, which_u, R_is_u):
    ob, gb = boxes
    cfg, _ = tiny
    R = cfg.ny + 1 if R_is_u else cfg.ny
    prod = rand_field(cfg, R, 12)
    vel = rand_field(cfg, R, 13)
    out_o = np.zeros_like(prod)
    ob.op("difvisxx", optr(prod), optr(vel), optr(out_o), R, which_u)
    dp, dv, do = DevArray.from_host(prod), DevArray.from_host(vel), DevArray(prod.size)
    gb.op("difvisxx", dp.ptr, dv.ptr, do.ptr, R, which_u)
    assert rel(do.to_host().reshape(prod.shape), out_o) < OP_TOL


@pytest.mark.parametrize("loguv", [0, 1])
def test_interpyy(boxes, tiny, loguv):
    ob, gb = boxes
    cfg, _ = tiny
    Rin = cfg.ny + 1 if loguv == 0 else cfg.ny
    g = rand_field(cfg, Rin, 14 + loguv)
    out_o = np.zeros((cfg.nx, cfg.ny + 1, cfg.nz1))
    ob.op("interpyy", optr(g), optr(out_o), loguv, cfg.nx)
    dg, do = DevArray.from_host(g), DevArray(out_o.size)
    gb.op("interpyy", dg.ptr, do.ptr, loguv, cfg.nx)
    out_g = do.to_host().reshape(out_o.shape)
    assert rel(out_g[:, :cfg.ny], out_o[:, :cfg.ny]) < OP_TOL


@pytest.mark.parametrize("which", [0, 1])
def test_differzy(boxes, tiny, which):
    ob, gb = boxes
    cfg, _ = tiny
    g = rand_field(cfg, cfg.ny + 1, 16 + which)
    out_o = np.zeros_like(g)
    ob.op("differzy", optr(g), optr(out_o), which, cfg.nx)
    dg, do = DevArray.from_host(g), DevArray(g.size)
    gb.op("differzy", dg.ptr, do.ptr, which, cfg.nx)
    rows = cfg.ny + 1 if which == 0 else cfg.ny
    out_g = do.to_host().reshape(g.shape)
    assert rel(out_g[:, :rows], out_o[:, :rows]) < OP_TOL


@pytest.mark.parametrize("R_is_u", [True, False])
def test_vistyy(boxes, tiny, R_is_u):
    ob, gb = boxes
    cfg, _ = tiny
    R = cfg.ny + 1 if R_is_u else cfg.ny
    u = rand_field(cfg, R, 18)
    out_o = np.zeros_like(u)
    ob.op("vistyy", optr(u), optr(out_o), R, cfg.nx)
    du, do = DevArray.from_host(u), DevArray(u.size)
    gb.op("vistyy", du.ptr, do.ptr, R, cfg.nx)
    assert rel(do.to_host().reshape(u.shape), out_o) < OP_TOL


@pytest.mark.parametrize("R_is_u", [True, False])
def test_implzy(boxes, tiny, R_is_u):
    ob, gb = boxes
    cfg, _ = tiny
    R = cfg.ny + 1 if R_is_u else cfg.ny
    u = rand_field(cfg, R, 19)
    res = rand_field(cfg, R, 20)
    rkk = -0.013
    u_o = u.copy()
    ob.op("implzy", optr(u_o), optr(res), R, rkk, cfg.nx)
    du, dr = DevArray.from_host(u), DevArray.from_host(res)
    gb.op("implzy", du.ptr, dr.ptr, R, rkk, cfg.nx)
    assert rel(du.to_host().reshape(u.shape), u_o) < OP_TOL


def test_fourxz(boxes, tiny, oracle_lib):
    ob, gb = boxes
    cfg, _ = tiny
    rng = np.random.default_rng(21)
    nrows = 37
    fou = rng.standard_normal((nrows, cfg.nz1))
    phys_o = np.zeros((nrows, cfg.nz))
    L = oracle_lib.lib()
    for r in range(nrows):
        L.oc_c2r(ob.h, optr(np.ascontiguousarray(fou[r])), optr(phys_o[r]))
    dfou, dphys = DevArray.from_host(fou), DevArray(nrows * cfg.nz)
    gb.op("fourxz", dfou.ptr, dphys.ptr, 1, nrows)
    assert rel(dphys.to_host().reshape(nrows, cfg.nz), phys_o) < OP_TOL
    phys = rng.standard_normal((nrows, cfg.nz))
    fou_o = np.zeros((nrows, cfg.nz1))
    for r in range(nrows):
        L.oc_r2c(ob.h, optr(np.ascontiguousarray(phys[r])), optr(fou_o[r]))
    dphys2, dfou2 = DevArray.from_host(phys), DevArray(nrows * cfg.nz1)
    gb.op("fourxz", dfou2.ptr, dphys2.ptr, -1, nrows)
    assert rel(dfou2.to_host().reshape(nrows, cfg.nz1), fou_o) < OP_TOL


@pytest.mark.parametrize("iopt", [-1, 1])
def test_cosftx(boxes, tiny, oracle_lib, iopt):
    ob, gb = boxes
    cfg, _ = tiny
    f = rand_field(cfg, cfg.ny, 22)
    lines = np.ascontiguousarray(f[1:, :cfg.ny - 1, :].transpose(1, 2, 0))
    oracle_lib.lib().oc_cosft(cfg.nx - 1, lines.shape[0] * lines.shape[1], optr(lines), iopt)
    din, dout = DevArray.from_host(f), DevArray(f.size)
    gb.op("cosftx", din.ptr, dout.ptr, iopt)
    out = dout.to_host().reshape(f.shape)[1:, :cfg.ny - 1, :].transpose(1, 2, 0)
    assert rel(out, lines) < OP_TOL


def _compare_state(ob, gb, tol, names=("u", "v", "w", "p")):
    errs = state_errors(gb, ob, names)
    assert_state(errs, tol)
    return errs


def test_rhsp_substep1(tiny, oracle_lib):
    cfg, st = tiny
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.rhsp(1)
    gb.rhsp(1)
    assert abs(gb.time()[1] - ob.time()[1]) <= 1e-14 * ob.time()[1]
    _compare_state(ob, gb, 1e-11, ("u", "v", "w", "p", "rhsupa", "rhsvpa", "rhswpa"))
    go, gg = ob.genflu_info(), gb.genflu_info()
    for k in go:
        assert abs(go[k] - gg[k]) <= 1e-11 * max(1.0, abs(go[k])), (k, go[k], gg[k])
    assert rel(gb.um(), ob.um()) < 1e-13
    eo, eg = ob.ener(), gb.ener()
    assert rel(eg, eo) < 1e-11, (eg, eo)
    gb.close()


def test_outflow_and_pois(tiny, oracle_lib):
    """outflow_correction and pois on a state with real divergence: seeded
    noise in modes 0..3 on interior rows (the generated state itself is
    divergence-free, which would leave p at round-off)."""
    cfg, st = tiny
    st = dict(st)
    rng = np.random.default_rng(5)
    for name in ("u", "v", "w"):
        a = st[name].copy()
        shp = a[1:, 1:-1, :4].shape
        a[1:, 1:-1, :4] += 1e-2 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
        a[:, :, 0] = a[:, :, 0].real
        st[name] = a
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    ob.outflow_correction()
    gb.outflow_correction()
    _compare_state(ob, gb, 1e-13, ("u", "rhsupa"))
    ob.pois(3.7)
    gb.pois(3.7)
    errs = _compare_state(ob, gb, 1e-11)
    print("pois relative errors:", errs)
    gb.close()


@pytest.mark.parametrize("nsteps", [1, 3])
def test_full_steps(tiny, oracle_lib, nsteps):
    cfg, st = tiny
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    for _ in range(nsteps):
        ob.step()
    gb.step(nsteps)
    errs = _compare_state(ob, gb, STEP_TOL)
    to, tg = ob.time(), gb.time()
    assert abs(to[0] - tg[0]) <= 1e-12 * to[0] and abs(to[1] - tg[1]) <= 1e-12 * to[1]
    print("relative errors after", nsteps, "steps:", errs)
    gb.close()


def test_frozen_dt_and_graphs(tiny, oracle_lib):
    cfg, st = tiny
    ob = oracle_lib.OracleBox(cfg, st, freeze_dt=True)
    ge = GpuBox(cfg, st, dt_mode=OTBL_DT_FROZEN)
    gg = GpuBox(cfg, st, dt_mode=OTBL_DT_FROZEN, graphs=True)
    for _ in range(2):
        ob.step()
    ge.step(2)
    # graph replays, with eager host work and syncs between them
    for _ in range(2):
        gg.step(1)
        gg.sync()
        _ = gg.time()
    _compare_state(ob, ge, STEP_TOL)
    for n in ("u", "v", "w", "p"):
        assert rel(gg.get(n), ge.get(n)) < 1e-14, n
    assert ob.time()[1] == pytest.approx(ge.time()[1], rel=1e-13)
    ge.close()
    gg.close()


def test_divergence_after_pois(tiny):
    """After pois the discrete divergence of pois.f90:41-54 is at round-off
    on planes 2..nx (the outflow correction zeroes the net flux first)."""
    cfg, st = tiny
    gb = GpuBox(cfg, st)
    gb.step(1)
    u, v, w = gb.get("u"), gb.get("v"), gb.get("w")
    kaz = ic.kaz_values(cfg)
    dy = np.diff(st["y"])
    h0 = cfg.ax * np.pi / (cfg.nx + 1)
    idx = 1.0 / (2.0 * h0 - 1.0 * h0)   # 1/dx(1), coeft.f90:401-420
    ny = cfg.ny
    div = (w[1:, 1:ny, :] * 1j * kaz[None, None, :]
           + (v[1:, 1:ny, :] - v[1:, 0:ny - 1, :]) / dy[1:ny][None, :, None]
           + idx * (u[1:, 1:ny, :] - u[:-1, 1:ny, :]))
    scale = np.abs(u).max() * idx
    assert np.abs(div).max() / scale < 1e-11
    gb.close()


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
@pytest.mark.parametrize("ny", [40, 97], ids=["y40", "y97"])
def test_fused_tail_matches_three_passes(mixed, ny, monkeypatch):
    """rhsp's tail fused per component (plane loop 2's y terms, boun's rows
    and rk_final in one pass; the default where the y solves run on
    full-wave chunks: FP64, and FP32 without the half-wave tables) gives the
    fields, rhs and rhs_prev of rhs_loop2 + boun + rk_final run apart
    (OTBL_Y_NOTAIL), bit for bit: the same chunks and the same expressions.
    ny = 40 gives chunks of 4 rows, ny = 97 of 8 (rows 2..4, from which boun
    extrapolates row 1, are exchanged through LDS either way)."""
    from opentbl_amd.lib import OTBL_MIXED
    cfg = config.small(name="tail", nx=130, ny=ny, nz=256, Lx=8.0 * 130 / 32, kzmax=8)
    st = ic.make_state(cfg)
    prec = {"precision": OTBL_MIXED} if mixed else {}
    if mixed:
        monkeypatch.setenv("OTBL_Y_HALF", "0") # the full-wave FP32 y kernels
    fused = GpuBox(cfg, st, graphs=True, **prec)
    monkeypatch.setenv("OTBL_Y_NOTAIL", "1")
    apart = GpuBox(cfg, st, graphs=True, **prec)
    assert fused.paths()["tail"] == "1" and apart.paths()["tail"] == "0"
    fused.step(3)
    apart.step(3)
    for n in ("u", "v", "w", "p", "rhsupa", "rhsvpa", "rhswpa"):
        a, b = fused.get(n), apart.get(n)
        assert np.isfinite(a).all(), n
        assert np.array_equal(a, b), (n, np.abs(a - b).max())
    fused.close()
    apart.close()


def test_half_wave_mixed_keeps_three_passes():
    """The mixed substep on half-wave y chunks keeps rhs_loop2, boun and
    rk_final apart (the fused kernel spilled there and was slower)."""
    from opentbl_amd.lib import OTBL_MIXED
    cfg = config.small(name="tailh", nx=130, ny=40, nz=256, Lx=8.0 * 130 / 32, kzmax=8)
    b = GpuBox(cfg, ic.make_state(cfg), precision=OTBL_MIXED)
    p = b.paths()
    assert p["y"].endswith("h") and p["tail"] == "0", p
    b.close()


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_side_stream_matches_serial(mixed, monkeypatch):
    """rhsp's plane loop 2 on its own stream beside plane loop 1 (the
    default where ya / yb hold no scratch) gives the fields the serial
    substep gives, bit for bit, in captured graphs replayed with host work
    between them and eagerly: the same kernels on the same inputs, so any
    missed dependency between the streams would show as a difference."""
    from opentbl_amd.lib import OTBL_MIXED
    cfg = config.small(name="side", nx=130, ny=40, nz=256, Lx=8.0 * 130 / 32, kzmax=8)
    st = ic.make_state(cfg)
    prec = {"precision": OTBL_MIXED} if mixed else {}
    two_g = GpuBox(cfg, st, graphs=True, **prec)
    two_e = GpuBox(cfg, st, **prec)
    monkeypatch.setenv("OTBL_SERIAL_LOOPS", "1")
    one = GpuBox(cfg, st, graphs=True, **prec)
    assert two_g.paths()["streams"] == "2" and one.paths()["streams"] == "1"
    for _ in range(3):
        two_g.step(1)
        two_g.sync()
        _ = two_g.time()
    two_e.step(3)
    one.step(3)
    for n in ("u", "v", "w", "p"):
        a, b, c = two_g.get(n), two_e.get(n), one.get(n)
        assert np.array_equal(a, c), n
        assert np.array_equal(b, c), n
    for b in (two_g, two_e, one):
        b.close()
