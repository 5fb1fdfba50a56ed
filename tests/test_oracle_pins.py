"""Pins of the CPU FP64 oracle against mathematics that fixes each part
(there are no golden vectors in the reference: it has no tests and no data
beyond its parameter files).  Every check below would fail for a plausible
transcription mistake: a wrong sign, index or term in a stencil, a dropped
boundary row, a transposed operand.
"""
import math

import numpy as np
import pytest
import scipy.fft as sfft

from opentbl_amd import config, ic


@pytest.fixture(scope="module")
def box(oracle_lib):
    cfg = config.small(nx=48, ny=40, nz=24)
    st = ic.make_state(cfg)
    return cfg, st, oracle_lib.OracleBox(cfg, st)


def fac_tridiag(c):
    """Rebuild the unfactored tridiagonal (l, d, u) from coefinv's LU
    storage (l = -a/d_prev, 1/d, c) -- coeft.f90:1338-1353."""
    n = c.shape[1]
    A = np.zeros((n, n))
    d = 1.0 / c[1]
    for i in range(n):
        A[i, i] = d[i] if i == 0 else d[i] + (-c[0, i]) * c[2, i - 1]
        if i > 0:
            A[i, i - 1] = -c[0, i] * d[i - 1]
        if i < n - 1:
            A[i, i + 1] = c[2, i]
    return A


# ---------------------------------------------------------------- grids
def x_grid(cfg):
    h0 = cfg.ax * math.pi / (cfg.nx + 1)
    return np.arange(cfg.nx + 2) * h0


def test_lele_interior_values(box):
    """Uniform-x interior rows reduce to the textbook 6th/4th-order Lele
    schemes: first derivative alpha = 1/22 ... is not used here; the
    staggered midpoint interpolation has alpha = 3/10 (coeft.f90:209-216 for
    the collocated version); on the staggered uniform grid the reference's
    non-uniform formula gives alpha = 1/6, a/2 = ... -- check the closed form
    of the implicit rows."""
    cfg, st, ob = box
    inxv = ob.coef("inxv")
    # unfactored implicit row i of inxv on a uniform grid: a^2/((b+a)(a+2b)) = 1/6
    A = fac_tridiag(inxv)
    i = cfg.nx // 2
    assert A[i, i - 1] == pytest.approx(1.0 / 6.0, rel=1e-12)
    assert A[i, i + 1] == pytest.approx(1.0 / 6.0, rel=1e-12)
    # explicit part (cofivx): 4a^2/((b+a)(2a+b)) = 2/3 on a uniform grid
    c = ob.coef("cofivx")
    assert c[0, i] == pytest.approx(2.0 / 3.0, rel=1e-12)
    assert c[1, i] == pytest.approx(2.0 / 3.0, rel=1e-12)
    # staggered first derivative on a uniform grid: alpha = 1/22, a = 12/11 h^-1
    A = fac_tridiag(ob.coef("dcxu"))
    assert A[i, i - 1] == pytest.approx(1.0 / 22.0, rel=1e-12)
    h = x_grid(cfg)[1]
    cc = ob.coef("cofcxu")
    assert cc[1, i] * h == pytest.approx(24.0 / 22.0, rel=1e-12)
    # second derivative: alpha = 1/10, explicit 6/5 h^-2
    A = fac_tridiag(ob.coef("vixu"))
    assert A[i, i - 1] == pytest.approx(1.0 / 10.0, rel=1e-12)
    ev = ob.coef("cofvxu")
    assert ev[0, i] * h * h == pytest.approx(12.0 / 10.0, rel=1e-12)
    assert ev[1, i] * h * h == pytest.approx(-12.0 / 5.0, rel=1e-12)


def _apply_interp_x(ob, cfg, f_u, which_u):
    """interpxx on one line through the operator entry point."""
    from oracle.oracle import ptr
    R = cfg.ny + 1
    g = np.zeros((cfg.nx, R, cfg.nz1))
    g[:, 0, 0] = f_u
    out = np.zeros_like(g)
    ob.op("interpxx", ptr(g), ptr(out), R, which_u)
    return out[:, 0, 0]


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_interpxx_polynomial_exactness(box, deg):
    """u -> midpoint interpolation (interpxx with the u coefficients) is
    exact on quadratics everywhere and on cubics away from the outflow end,
    whose 3-point closure (inbx row 2, coefb flag 0) is second order; its
    error decays geometrically away from the last row."""
    cfg, st, ob = box
    x = x_grid(cfg)
    xu = x[1:cfg.nx + 1]                    # u at x(i)
    xm = 0.5 * (x[1:cfg.nx + 1] + x[2:cfg.nx + 2])  # midpoints
    f = lambda t: (t / x[-1]) ** deg
    got = _apply_interp_x(ob, cfg, f(xu), 1)
    err = np.abs(got - f(xm))
    if deg <= 2:
        assert err.max() < 1e-11
    else:
        assert err[: cfg.nx // 2].max() < 1e-11
        assert err[-1] > 1e-7


def _y_apply(ob, cfg, name, col, extra=None):
    from oracle.oracle import ptr
    R = cfg.ny + 1
    g = np.zeros((cfg.nx, R, cfg.nz1))
    g[0, :len(col), 0] = col
    out = np.zeros_like(g)
    if name == "interpyy":
        ob.op("interpyy", ptr(g if extra == 0 else np.ascontiguousarray(g[:, :cfg.ny])),
              ptr(out), extra, cfg.nx)
    elif name == "differzy":
        ob.op("differzy", ptr(g), ptr(out), extra, cfg.nx)
    return out[0, :, 0]


@pytest.fixture(scope="module")
def ubox(oracle_lib):
    """A box on a uniform y grid, where the reference's non-uniform compact
    formulas reduce to the textbook staggered schemes."""
    cfg = config.small(nx=32, ny=40, nz=24, stretch=0.0)
    st = ic.make_state(cfg, with_perturbation=False)
    return cfg, st, oracle_lib.OracleBox(cfg, st)


def _ycols(st, cfg):
    y = st["y"]
    return 0.5 * (y[:-1] + y[1:]), y[1:cfg.ny + 1]


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_interpyy_u_to_v_exact_uniform(ubox, deg):
    """interpyy u rows -> v rows (lip.f90:376-392): exact for cubics that
    vanish at the wall (row 1 is set to zero for no-slip)."""
    cfg, st, ob = ubox
    yu, yv = _ycols(st, cfg)
    f = lambda t: (t / 4.0) ** deg
    got = _y_apply(ob, cfg, "interpyy", f(yu), 0)[:cfg.ny]
    assert np.abs(got - f(yv)).max() < 1e-12
    assert got[0] == 0.0


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_interpyy_v_to_u_exact_uniform(ubox, deg):
    """interpyy v rows -> u rows 1..ny (lip.f90:393-410) with its one-sided
    end rows inby(3,:), inby(4,:)."""
    cfg, st, ob = ubox
    yu, yv = _ycols(st, cfg)
    f = lambda t: ((t + 0.5) / 4.0) ** deg
    got = _y_apply(ob, cfg, "interpyy", f(yv), 1)[:cfg.ny]
    tol = 1e-12 if deg <= 2 else 1e-4   # the 3-point closure inby(3,:) is second order
    assert np.abs(got - f(yu[:cfg.ny])).max() < tol
    assert np.abs(got[cfg.ny // 2:] - f(yu[cfg.ny // 2:cfg.ny])).max() < 1e-11


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_differzy_exact_uniform(ubox, deg):
    """d/dy v rows -> u rows (dcyu) and u rows -> v rows (dcyv): the
    staggered compact first derivative and its 4-point closures are exact
    for cubics on a uniform grid."""
    cfg, st, ob = ubox
    yu, yv = _ycols(st, cfg)
    f = lambda t: ((t + 0.5) / 4.0) ** deg
    df = lambda t: deg * ((t + 0.5) / 4.0) ** (deg - 1) / 4.0
    col = np.zeros(cfg.ny + 1)
    col[:cfg.ny] = f(yv)
    got = _y_apply(ob, cfg, "differzy", col, 0)[:cfg.ny + 1]
    assert np.abs(got - df(yu)).max() < 1e-11
    col = f(yu).copy()
    got = _y_apply(ob, cfg, "differzy", col, 1)[:cfg.ny]
    assert np.abs(got - df(yv)).max() < 1e-11


def test_vistyy_exact_uniform(ubox):
    """vistyy: the compact second derivative (interior alpha = 1/10) and its
    closures (coefb flag 1) are exact for polynomials up to degree 3."""
    from oracle.oracle import ptr
    cfg, st, ob = ubox
    yu, yv = _ycols(st, cfg)
    for R, yy in ((cfg.ny + 1, yu), (cfg.ny, yv)):
        for deg in (0, 1, 2, 3):
            f = ((yy + 0.5) / 4.0) ** deg
            d2 = deg * (deg - 1) * ((yy + 0.5) / 4.0) ** max(deg - 2, 0) / 16.0
            g = np.zeros((cfg.nx, R, cfg.nz1))
            g[0, :, 0] = f
            out = np.zeros_like(g)
            ob.op("vistyy", ptr(g), ptr(out), R, cfg.nx)
            assert np.abs(out[0, :, 0] - d2).max() < 1e-9, (R, deg)


def _stretched_err(oracle_lib, ny, which):
    """Max error of an operator on a smooth function on the stretched grid."""
    from oracle.oracle import ptr
    cfg = config.small(nx=16, ny=ny, nz=12, kzmax=2)
    st = ic.make_state(cfg, with_perturbation=False)
    ob = oracle_lib.OracleBox(cfg, st)
    yu, yv = _ycols(st, cfg)
    f = lambda t: np.sin(0.9 * t) * t
    df = lambda t: 0.9 * np.cos(0.9 * t) * t + np.sin(0.9 * t)
    d2f = lambda t: -0.81 * np.sin(0.9 * t) * t + 1.8 * np.cos(0.9 * t)
    R = cfg.ny + 1
    g = np.zeros((cfg.nx, R, cfg.nz1))
    out = np.zeros_like(g)
    if which == "differzy":
        g[0, :cfg.ny, 0] = f(yv)
        ob.op("differzy", ptr(g), ptr(out), 0, cfg.nx)
        return np.abs(out[0, :, 0] - df(yu)).max()
    if which == "interpyy":
        g[0, :, 0] = f(yu)
        ob.op("interpyy", ptr(g), ptr(out), 0, cfg.nx)
        return np.abs(out[0, :cfg.ny, 0] - f(yv)).max()
    g[0, :, 0] = f(yu)
    ob.op("vistyy", ptr(g), ptr(out), R, cfg.nx)
    return np.abs(out[0, 1:-1, 0] - d2f(yu[1:-1])).max()


@pytest.mark.parametrize("which", ["differzy", "interpyy", "vistyy"])
def test_stretched_grid_convergence(oracle_lib, which):
    """On the tanh-stretched grid the non-uniform schemes are consistent:
    doubling ny cuts the error of a smooth function by at least 4x."""
    e1 = _stretched_err(oracle_lib, 40, which)
    e2 = _stretched_err(oracle_lib, 80, which)
    assert e1 < 1e-1 and e2 < e1 / 4.0, (e1, e2)


def test_tridiagonal_sweeps_match_dense_solve(box):
    """The LU sweep of interpxx equals a dense solve of the assembled
    tridiagonal system with the explicit right-hand side."""
    cfg, st, ob = box
    inxu = ob.coef("inxu")
    A = fac_tridiag(inxu)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(cfg.nx)
    e = ob.coef("cofiux")
    inbx = ob.coef("inbx")
    b = np.zeros(cfg.nx)
    b[1:-1] = e[0, 1:-1] * f[1:-1] + e[1, 1:-1] * f[2:]
    b[0] = inbx[2, :] @ f[:4]                 # chi2 = chi+1 = 3 for u
    b[-1] = inbx[1, :] @ f[::-1][:4]          # chi = 2
    ref = np.linalg.solve(A, b)
    got = _apply_interp_x(ob, cfg, f, 1)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_kmod_closed_form(box):
    """kmod(i) are the eigenvalues of the Neumann second difference
    (coeft.f90:1181-1183): check against a dense eigen-decomposition."""
    cfg, st, ob = box
    km = ob.coef("kmod")
    n = cfg.nx - 1
    h = x_grid(cfg)[2] - x_grid(cfg)[1]
    L = (np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))
    L[0, 0] = L[-1, -1] = -1.0
    ev = np.sort(-np.linalg.eigvalsh(L)) / h ** 2
    assert np.abs(np.sort(km) - ev).max() < 1e-9 * ev.max()


def test_mtr_is_div_of_grad(box):
    """ayp (mtr, coeft.f90:1244-1332) equals the y part of div(grad) as pois
    uses them: div with idy(j) on v differences, grad with idyy(j)."""
    cfg, st, ob = box
    ayp = ob.coef("ayp")
    ny = cfg.ny
    dy = np.diff(st["y"])
    idy = 1.0 / dy
    idyy = 2.0 / (dy[:-1] + dy[1:])      # idyy(1:ny)
    # phi at p rows j = 1..ny-1 (rest rows); v_j = (phi_j - phi_{j-1}) idyy(j) at v rows 2..ny-1
    n = ny - 1
    G = np.zeros((ny, n))
    for j in range(2, ny):               # v rows 2..ny-1 (1-based)
        G[j - 1, j - 1] = idyy[j - 1]
        G[j - 1, j - 2] = -idyy[j - 1]
    D = np.zeros((n, ny))
    for j in range(1, ny):               # div rows 1..ny-1: (v(j+1)-v(j)) idy(j)
        D[j - 1, j] = idy[j]
        D[j - 1, j - 1] = -idy[j]
    L = D @ G
    A = np.zeros((n, n))
    for j in range(n):
        A[j, j] = ayp[1, j]
        if j > 0:
            A[j, j - 1] = ayp[0, j]
        if j < n - 1:
            A[j, j + 1] = ayp[2, j]
    assert np.abs(A - L).max() < 1e-9 * np.abs(L).max()


def test_fourxz_matches_numpy(box, oracle_lib):
    """c2r / r2c with the reference scaling (ffts.f90:48-58): unnormalised
    inverse with modes above nz2 zeroed, forward times 1/nz."""
    from oracle.oracle import ptr
    cfg, st, ob = box
    rng = np.random.default_rng(1)
    X = rng.standard_normal(cfg.nz1)
    phys = np.zeros(cfg.nz)
    oracle_lib.lib().oc_c2r(ob.h, ptr(X), ptr(phys))
    spec = np.zeros(cfg.nz // 2 + 1, complex)
    spec[:cfg.mz] = X[0::2] + 1j * X[1::2]
    spec[0] = spec[0].real
    ref = np.fft.irfft(spec, cfg.nz) * cfg.nz
    assert np.abs(phys - ref).max() < 1e-13 * np.abs(ref).max()
    back = np.zeros(cfg.nz1)
    oracle_lib.lib().oc_r2c(ob.h, ptr(ref.copy()), ptr(back))
    ref2 = np.fft.rfft(ref)[:cfg.mz] / cfg.nz
    got = back[0::2] + 1j * back[1::2]
    assert np.abs(got - ref2).max() < 1e-13
    assert back[1] == 0.0


@pytest.mark.parametrize("n", [2, 3, 10, 31, 255, 383, 1023, 1535, 1536, 3071])
def test_cosft_matches_scipy(oracle_lib, n):
    """cosftx (ffts.f90:147-181): FFTW REDFT10 / n and REDFT01 * 0.5 are
    scipy's unnormalised DCT-II / n and DCT-III * 0.5; round trip = 1.
    The lengths cover the oracle's mixed-radix path (every prime factor at
    most 64: 255, 1023, 1536) and its Bluestein path (a prime factor above
    64: 383, 1535 = 5 * 307, 3071 = 37 * 83)."""
    from oracle.oracle import ptr
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n))
    a = x.copy()
    oracle_lib.lib().oc_cosft(n, 3, ptr(a), -1)
    assert np.abs(a - sfft.dct(x, 2, axis=1) / n).max() < 1e-13
    c = x.copy()
    oracle_lib.lib().oc_cosft(n, 3, ptr(c), 1)
    assert np.abs(c - 0.5 * sfft.dct(x, 3, axis=1)).max() < 1e-12 * np.abs(c).max()
    b = a.copy()
    oracle_lib.lib().oc_cosft(n, 3, ptr(b), 1)
    assert np.abs(b - x).max() < 1e-12


def _div(cfg, st_or_box, u, v, w):
    kaz = ic.kaz_values(cfg)
    dy = np.diff(st_or_box["y"])
    h0 = cfg.ax * np.pi / (cfg.nx + 1)
    idx = 1.0 / (2.0 * h0 - 1.0 * h0)
    ny = cfg.ny
    return (w[1:, 1:ny, :] * 1j * kaz[None, None, :]
            + (v[1:, 1:ny, :] - v[1:, 0:ny - 1, :]) / dy[1:ny][None, :, None]
            + idx * (u[1:, 1:ny, :] - u[:-1, 1:ny, :]))


def test_pois_projects_to_zero_divergence(oracle_lib):
    """After outflow_correction and pois (pois.f90:11-189) the discrete
    divergence is at round-off on planes 2..nx, for a state with O(1e-2)
    divergence."""
    cfg = config.small()
    st = dict(ic.make_state(cfg))
    rng = np.random.default_rng(9)
    for name in ("u", "v", "w"):
        a = st[name].copy()
        shp = a[1:, 1:-1, :3].shape
        a[1:, 1:-1, :3] += 1e-2 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
        a[:, :, 0] = a[:, :, 0].real
        st[name] = a
    ob = oracle_lib.OracleBox(cfg, st)
    d0 = np.abs(_div(cfg, st, ob.get("u"), ob.get("v"), ob.get("w"))).max()
    ob.outflow_correction()
    ob.pois(1.0)
    d1 = np.abs(_div(cfg, st, ob.get("u"), ob.get("v"), ob.get("w"))).max()
    assert d0 > 1e-2 and d1 < 1e-12 * d0 * 1e3


def test_outflow_correction_zeroes_net_flux(oracle_lib):
    """The mass-flux correction (boun.f90:63-93) makes the net flux of the
    mean mode through inflow, outflow and top exactly zero (to round-off)."""
    cfg = config.small()
    st = dict(ic.make_state(cfg))
    st["u"] = st["u"].copy()
    st["u"][-1, 1:cfg.ny, 0] *= 1.01
    ob = oracle_lib.OracleBox(cfg, st)
    ob.outflow_correction()
    u, v = ob.get("u"), ob.get("v")
    dy = np.diff(st["y"])
    h0 = cfg.ax * np.pi / (cfg.nx + 1)
    dx1 = 2.0 * h0 - 1.0 * h0
    net = (v[1:, cfg.ny - 1, 0].real.sum() * dx1
           - (u[0, 1:cfg.ny, 0].real * dy[1:cfg.ny]).sum()
           + (u[-1, 1:cfg.ny, 0].real * dy[1:cfg.ny]).sum())
    assert abs(net) < 1e-13


def test_boundary_rows_satisfy_their_identities(oracle_lib):
    """After rhsp (boun, boun.f90:7-56, and the right-hand-side rows,
    rhsp.F90:423-445): v = 0 at the wall, the top mean v equals vmagic
    (implzy keeps rows 1 and ny), the saved rhs is zero on plane 1, on v rows
    1 and ny and on u, w row ny+1, and its row 1 satisfies the wall
    extrapolation inby(2,:)."""
    cfg = config.small()
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    ob.rhsp(1)
    v = ob.get("v")
    inby, vm = ob.coef("inby"), ob.coef("vmagic")
    assert np.abs(v[:, 0]).max() == 0.0
    assert np.abs(v[:, -1, 0].real - vm).max() == 0.0
    assert np.abs(v[:, -1, 1:]).max() == 0.0
    ru, rv, rw = ob.get("rhsupa"), ob.get("rhsvpa"), ob.get("rhswpa")
    assert np.abs(ru[0]).max() == 0 and np.abs(rv[0]).max() == 0 and np.abs(rw[0]).max() == 0
    assert np.abs(rv[:, 0]).max() == 0 and np.abs(rv[:, -1]).max() == 0
    for a in (ru, rw):
        assert np.abs(a[:, -1]).max() == 0
        r1 = inby[1, 0] * a[:, 0] + inby[1, 1] * a[:, 1] + inby[1, 2] * a[:, 2] + inby[1, 3] * a[:, 3]
        assert np.abs(r1).max() < 1e-13 * max(1.0, np.abs(a).max())


def test_genflu_identity_rescaling(oracle_lib):
    """genflu with Re_theta,out = Re_theta,in (gamma = 1): the rescaled
    inflow modes >= 1 equal the blended recycled plane (the interpolation
    weights reduce to the identity), the mode-0 plane is u0/v0."""
    cfg = config.small()
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    u_before = ob.get("u")
    ob.rhsp(1)
    info = ob.genflu_info()
    assert info["gamma2"] == pytest.approx(1.0, abs=1e-3)
    u = ob.get("u")
    # rows 2..ny of the mean mode are u0 (boun then rewrites the ghost rows)
    assert np.abs(u[0, 1:cfg.ny, 0].real - st["u0"][1:cfg.ny]).max() == 0.0
    # with gamma ~ 1 the rescaled plane stays close to the blended copy of plane xout
    y = st["y"]
    jy = min(119, cfg.ny)
    fb = 0.5 * (1 - np.tanh(y[1:cfg.ny + 2] - y[jy]))
    ref = u_before[cfg.xout - 1, :, 1:] * fb[:, None]
    assert np.abs(u[0, 3:-3, 1:] - ref[3:-3]).max() < 5e-2 * np.abs(ref).max()


def test_viscous_decay_of_a_spanwise_mode(oracle_lib):
    """A single spanwise mode on a uniform-in-x, y-independent interior field
    decays by the explicit z-viscous term of the RK3 step: after one
    substep the mode changes by ~ -gamma dt nu kz^2 u (rhsp.F90:355-365 and
    the combination rhsp.F90:448-475)."""
    cfg = config.small(kzmax=0)
    st = dict(ic.make_state(cfg, with_perturbation=False))
    st["u"] = np.zeros_like(st["u"])
    st["v"] = np.zeros_like(st["v"])
    st["w"] = np.zeros_like(st["w"])
    eps = 1e-6
    st["u"][:, :, 1] = eps
    ob = oracle_lib.OracleBox(cfg, st)
    ob.set("u", st["u"])
    ob.rhsp(1)
    dt = ob.time()[1]
    kz = ic.kaz_values(cfg)[1]
    u = ob.get("u")
    i, j = cfg.nx // 2, cfg.ny // 2
    expect = eps * (1.0 - (8.0 / 15.0) * dt * kz ** 2 / st["re"])
    # the implicit y step leaves a y-constant interior unchanged
    assert u[i, j, 1].real == pytest.approx(expect, rel=1e-8)


def test_fp32_transposes_round_at_the_transposes(oracle_lib):
    """The reference rounds through FP32 at every plane<->pencil transpose
    (change.f90:565-686).  The oracle's emulation of that differs from the
    FP64-transpose variant at the FP32 level, not more."""
    cfg = config.small()
    st = ic.make_state(cfg)
    a = oracle_lib.OracleBox(cfg, st)
    b = oracle_lib.OracleBox(cfg, st, fp32_transposes=True)
    a.step()
    b.step()
    ua, ub = a.get("u"), b.get("u")
    rel = np.abs(ua - ub).max() / np.abs(ua).max()
    assert 1e-9 < rel < 1e-5


def test_info_per_step_pins(oracle_lib):
    """The numpy info_per_step (the diagnostics records' per-plane columns)
    against the C oracle's inlet_retheta (bl_1.F90:372-396: the delta_99 and
    Re_theta formulas of info_per_step, restated separately) on a field whose
    mean profile is the inflow profile u0 at every plane, and against closed
    forms: the sample-plane indices, u_inf of a flat top, u_tau of a linear
    wall profile, delta_99 of a ramp."""
    from oracle.oracle import info_per_step_planes
    cfg = config.small()
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    nx, ny, y = cfg.nx, cfg.ny, np.asarray(st["y"])
    u = np.zeros((nx, ny + 1, 2), complex)
    u[:, :, 0] = st["u0"]
    out = info_per_step_planes(u, y, st["re"])
    assert list(out[0]) == [((nx - 1) * i) // 8 + 1 for i in range(1, 9)]
    assert np.all(out[3] == ob.scalar("din"))
    assert np.abs(out[4] - ob.scalar("rthin")).max() <= 1e-13 * ob.scalar("rthin")
    # ramp U min(1, (y - y1)/delta): slope U/delta at the wall, flat top
    U, delta = 1.3, 0.5 * (y[ny // 2] - y[1])
    prof = U * np.minimum(1.0, (y[1:ny + 2] - y[1]) / delta)
    u[:, :, 0] = prof
    out = info_per_step_planes(u, y, st["re"])
    assert np.allclose(out[1], U, rtol=1e-15)
    assert np.allclose(out[2], np.sqrt(U / delta / st["re"]), rtol=1e-12)
    # the first row at or above 0.99 U (float32 0.99, the reference's literal)
    j99 = next(j for j in range(1, ny + 1) if prof[j - 1] >= float(np.float32(0.99)) * U)
    assert np.all(out[3] == y[j99])
    assert abs(y[j99] - (y[1] + 0.99 * delta)) <= (y[j99] - y[j99 - 1])
    # momentum thickness of the ramp, U^2 delta / 6: the reference's sum is a
    # first-order rectangle rule, so on a fine grid (ny = 320) within 2%
    cf = config.small(ny=320)
    sf = ic.make_state(cf)
    yf = np.asarray(sf["y"])
    df = 0.5 * (yf[cf.ny // 2] - yf[1])
    uf = np.zeros((cf.nx, cf.ny + 1, 1))
    uf[:, :, 0] = U * np.minimum(1.0, (yf[1:cf.ny + 2] - yf[1]) / df)
    theta = info_per_step_planes(uf, yf, sf["re"])[4] / sf["re"] / U
    assert np.abs(theta - df / 6).max() < 0.02 * df / 6


# ------------------------------------------------ difvisxx (lip.f90:52-142)
def _difvis(ob, cfg, prod, vel, which_u):
    from oracle.oracle import ptr
    R = cfg.ny + 1
    g = np.zeros((cfg.nx, R, cfg.nz1))
    q = np.zeros_like(g)
    g[:, 0, 0] = prod
    q[:, 0, 0] = vel
    out = np.zeros_like(g)
    ob.op("difvisxx", ptr(g), ptr(q), ptr(out), R, which_u)
    return out[:, 0, 0]


@pytest.mark.parametrize("which_u", [1, 0])
@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_difvisxx_convective_exact_on_cubics(box, which_u, deg):
    """The convective part of difvisxx (the compact staggered first
    derivative, tridiagonal dcx* with the 2-point right-hand side and the
    4-point boundary rows dcbx) is exact on cubics at every row but the
    outflow row: the product, given at the midpoints x(i + 1/2), is
    differentiated at x(i).  (The x grid is uniform, so the u and v/w
    coefficient sets coincide up to round-off: both are checked.)"""
    cfg, st, ob = box
    x = x_grid(cfg)
    h = x[1]
    xf = x[1:cfg.nx + 1]
    L = x[-1]
    got = _difvis(ob, cfg, ((xf + h / 2) / L) ** deg, np.zeros(cfg.nx), which_u)
    want = deg * (xf / L) ** max(deg - 1, 0) / L if deg else np.zeros(cfg.nx)
    assert np.abs(got[:-1] - want[:-1]).max() < 1e-11


@pytest.mark.parametrize("which_u", [1, 0])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_difvisxx_viscous_exact_on_quartics(box, which_u, deg):
    """The viscous part, rex d2/dx2 of the velocity (tridiagonal vix*, 3-point
    right-hand side, 4-point boundary rows cofvbx), is exact on quartics at
    every row but the outflow row, and degree 5 already shows the closures'
    truncation error."""
    cfg, st, ob = box
    x = x_grid(cfg)
    xf = x[1:cfg.nx + 1]
    L = x[-1]
    rex = -1.0 / st["re"]
    got = _difvis(ob, cfg, np.zeros(cfg.nx), (xf / L) ** deg, which_u)
    want = rex * deg * (deg - 1) * (xf / L) ** max(deg - 2, 0) / L ** 2
    # round-off scale: the stencil weights are of order rex / h^2
    scale = max(np.abs(want).max(), abs(rex) / x[1] ** 2)
    assert np.abs(got[:-1] - want[:-1]).max() < 1e-11 * scale
    got5 = _difvis(ob, cfg, np.zeros(cfg.nx), (xf / L) ** 5, which_u)
    want5 = rex * 20 * (xf / L) ** 3 / L ** 2
    assert np.abs(got5[:-1] - want5[:-1]).max() > 1e-9 * np.abs(want5).max()


@pytest.mark.parametrize("which_u", [1, 0])
def test_difvisxx_outflow_row(box, which_u):
    """At the outflow plane the convective derivative is replaced by the
    convective outflow condition U_inf (q(nx) - q(nx-1)) / dx with U_inf = 1,
    and the viscous term by zero (lip.f90:105,132), whatever the product."""
    cfg, st, ob = box
    rng = np.random.default_rng(5)
    prod, vel = rng.standard_normal(cfg.nx), rng.standard_normal(cfg.nx)
    h = x_grid(cfg)[1]
    got = _difvis(ob, cfg, prod, vel, which_u)
    assert got[-1] == pytest.approx((vel[-1] - vel[-2]) / h, rel=1e-14)


# ---------------------------------------------------- implzy (lip.f90:528-597)
@pytest.mark.parametrize("uu", [True, False], ids=["u_rows", "v_rows"])
def test_implzy_dense_solve(box, uu):
    """implzy solves (A + rkk B) u = A wki on rows 2..jee-1 with A the
    unfactored compact second-derivative matrix (vyui / vyvi, the copy made
    before coefinv) and B its explicit stencil (cofvyu / cofvyv, entries for
    rows j+1, j, j-1), rows 1 and jee of u kept (Dirichlet): against a dense
    solve of the assembled system."""
    from oracle.oracle import ptr
    cfg, st, ob = box
    jee = cfg.ny + 1 if uu else cfg.ny
    A = ob.coef("vyui" if uu else "vyvi")      # (3, jee): column j-1 is row j
    B = ob.coef("cofvyu" if uu else "cofvyv")  # (3, ny)
    rkk = -0.37
    rng = np.random.default_rng(9)
    u0 = rng.standard_normal((1, jee, cfg.nz1))
    wki = rng.standard_normal((1, jee, cfg.nz1))
    got = u0.copy()
    ob.op("implzy", ptr(got), ptr(wki), jee, rkk, 1)
    M = np.eye(jee)
    rhs = u0[0].copy()
    for j in range(2, jee):                   # 1-based rows 2..jee-1
        r = j - 1
        M[r, r - 1] = A[0, r] + rkk * B[2, r]
        M[r, r] = A[1, r] + rkk * B[1, r]
        M[r, r + 1] = A[2, r] + rkk * B[0, r]
        rhs[r] = A[0, r] * wki[0, r - 1] + A[1, r] * wki[0, r] + A[2, r] * wki[0, r + 1]
    ref = np.linalg.solve(M, rhs)
    assert np.abs(got[0] - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("uu", [True, False], ids=["u_rows", "v_rows"])
def test_implzy_inverts_the_explicit_viscous_operator(box, uu):
    """The implicit step is the inverse of I + rkk D_yy with D_yy the
    explicit compact second derivative vistyy (lip.f90:447-513) applies:
    for wki = u* + rkk D_yy u* (rows 2..jee-1) and the Dirichlet rows of u*,
    implzy returns u* (the implicit system is the explicit operator's own
    matrix pair, coeft.f90:1116-1134)."""
    from oracle.oracle import ptr
    cfg, st, ob = box
    jee = cfg.ny + 1 if uu else cfg.ny
    rkk = -0.05
    rng = np.random.default_rng(4)
    us = rng.standard_normal((1, jee, cfg.nz1))
    d = np.zeros_like(us)
    ob.op("vistyy", ptr(us), ptr(d), jee, 1)
    wki = us + rkk * d
    got = us.copy()
    got[0, 1:-1] = 0.0
    ob.op("implzy", ptr(got), ptr(wki), jee, rkk, 1)
    assert np.abs(got - us).max() < 1e-12 * np.abs(us).max()


# ------------------------------------------- Re_theta(x), Cf(x) diagnostics
def test_retheta_cf_closed_forms(oracle_lib):
    """Re_theta(x) and Cf(x) of the kz = 0 u mode (genflu.F90:76-87): equal
    to the inlet values of inlet_retheta (bl_1.F90:372-396, the same sums
    restated separately) where the plane's profile is u0; for a ramp
    U min(1, (y - y1) / delta) Cf = 2 U / (delta Re) exactly and Re_theta =
    Re U delta / 6 up to the sum's first-order quadrature."""
    cfg = config.small(ny=320)
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    y = np.asarray(st["y"])
    ny, re = cfg.ny, st["re"]
    u = ob.get("u")
    u[:, :, 0] = st["u0"]
    # planes 5.. take ramps of growing thickness
    U = 1.1
    deltas = np.linspace(0.3, 0.6, cfg.nx - 4)
    yu = 0.5 * (y[0:ny + 1] + y[1:ny + 2])
    for q, dl in enumerate(deltas):
        u[4 + q, :, 0] = U * np.minimum(1.0, (yu - y[1]) / dl)
    ob.set("u", u)
    rth, cf = ob.retheta_cf()
    assert np.abs(rth[:4] - ob.scalar("rthin")).max() <= 1e-13 * ob.scalar("rthin")
    assert np.allclose(cf[4:], 2.0 * U / (deltas * re), rtol=1e-12)
    assert np.abs(rth[4:] / (re * U * deltas / 6.0) - 1.0).max() < 0.02


# ------------------------------------------ genflu rescaling (genflu.F90:88-242)
def _dobleu_integral(eta):
    """W(eta_i) = integral from eta_1 to eta_i of the inner/outer weight
    0.5 (1 + coth(4) tanh(4 (s - 0.2) / (0.6 s + 0.2))) (1 for s >= 1) of
    dobleu (genflu.F90:175-203), by adaptive quadrature, and its value at
    eta = 1."""
    from scipy.integrate import quad

    def wgt(s):
        return 1.0 if s >= 1.0 else 0.5 * (1.0 + math.tanh(4.0 * (s - 0.2) / (0.6 * s + 0.2)) / math.tanh(4.0))
    W = np.zeros(len(eta))
    for i in range(1, len(eta)):
        W[i] = W[i - 1] + quad(wgt, eta[i - 1], eta[i], points=[1.0] if eta[i - 1] < 1.0 < eta[i] else None)[0]
    w1 = quad(wgt, eta[0], 1.0, points=[0.2])[0]
    return W, w1


def test_genflu_rescaling_far_from_identity(oracle_lib):
    """The recycled inflow of genflu with gamma far from 1: the outlet plane
    xout carries a mean profile twice as thick as the inflow profile u0 (so
    Re_theta,out ~ 2 Re_theta,in and gamma_2 ~ 0.5), and, in every mode >= 1,
    a profile that after the tanh blend of genflu.F90:55-60 is linear in the
    rescaled coordinate ysca, a + b ysca.  Linear interpolation reproduces a
    linear profile, so the inflow modes must be (a + b wlwa)/gamma_1 with
    wlwa = ysca/gamma_1 + (1/gamma_2 - 1/gamma_1) W(ysca)/W(1), rows past
    ntop - 3 copies of row ntop - 3, the u/w ghost row the negative of row 2,
    and mode 0 the inflow profiles.  gamma_2 = Re_theta,in / Re_theta,out
    with Re_theta,out the Re_theta of the outlet profile (the running mean um
    is set to it, so the update of genflu.F90:67-75 keeps it), gamma_1 =
    gamma_2^(1/8) exactly; W, the weight integral of dobleu, to its midpoint
    rule's accuracy against quadrature."""
    cfg = config.small(ny=64, nx=32)
    st = dict(ic.make_state(cfg))
    ny, s = cfg.ny, cfg.nz1
    y = np.asarray(st["y"])
    yu = 0.5 * (y[0:ny + 1] + y[1:ny + 2])       # u rows 1..ny+1
    u0 = np.asarray(st["u0"])
    uout = np.interp(yu / 2.0, yu, u0)          # twice as thick
    uout[0] = -uout[1]
    probe = oracle_lib.OracleBox(cfg, st)
    din, rthin = probe.scalar("din"), probe.scalar("rthin")
    jy = min(119, ny)
    fblend = 0.5 * (1.0 - np.tanh(y[1:ny + 2] - y[jy]))
    ys_u = 0.5 * (y[1:ny + 1] + y[2:ny + 2]) / din    # ysca of u, i = 1..ny (row i+1)
    ys_v = y[1:ny + 1] / din
    rng = np.random.default_rng(12)
    a = rng.standard_normal(s - 2) * 1e-2
    b = rng.standard_normal(s - 2) * 1e-2
    u, v, w = (np.array(st[n]) for n in ("u", "v", "w"))
    xo = cfg.xout - 1
    u[xo, :, 0] = uout
    for arr, fac in ((u, 1.0), (w, -0.5)):
        pr = np.zeros((ny + 1, s - 2))
        pr[1:] = (a[None, :] + b[None, :] * ys_u[:, None]) * fac / fblend[1:, None]
        arr.view(np.float64).reshape(cfg.nx, ny + 1, s)[xo, :, 2:] = pr
    v.view(np.float64).reshape(cfg.nx, ny, s)[xo, :, 2:] = 0.7 * (a[None, :] + b[None, :] * ys_v[:, None])
    st.update(u=u, v=v, w=w, um=uout.copy())
    ob = oracle_lib.OracleBox(cfg, st)
    rth, _ = ob.retheta_cf()
    ob.rhsp(1)
    info = ob.genflu_info()
    g2 = rthin / rth[xo]
    assert info["gamma2"] == pytest.approx(g2, rel=1e-13)
    assert 0.4 < g2 < 0.6
    g1 = info["gamma1"]
    assert g1 == pytest.approx(g2 ** 0.125, rel=1e-15)
    uo, vo, wo = (ob.get(n) for n in ("u", "v", "w"))
    # mode 0: the inflow profiles (rows the boundary conditions leave alone)
    assert np.abs(uo[0, 1:ny, 0].real - u0[1:ny]).max() == 0.0
    assert np.abs(vo[0, 1:ny - 1, 0].real - np.asarray(st["v0"])[1:ny - 1]).max() == 0.0
    for kind, ys, out, fac, first in (("u", ys_u, uo, 1.0, 1), ("w", ys_u, wo, -0.5, 1),
                                      ("v", ys_v, vo, 0.7, 0)):
        W, w1 = _dobleu_integral(ys)
        wl = ys / g1 + (1.0 / g2 - 1.0 / g1) / w1 * W
        ntop = max(i + 1 for i in range(ny) if wl[i] <= ys[-1])
        R = out.shape[1]
        got = out[0].view(np.float64).reshape(R, s)[:, 2:]
        want = np.zeros((R, s - 2))
        rows = np.arange(1, ntop + 1)                 # i = 1..ntop
        vals = fac * (a[None, :] + b[None, :] * wl[rows - 1][:, None]) / g1
        want[rows - 1 + first] = vals                 # u/w: row i+1, v: row i
        want[ntop - 3:] = want[ntop - 4]              # rows ntop-2.. copy row ntop-3
        # rows the boundary conditions rewrite afterwards are left out
        lo, hi = (1, R - 1) if kind != "v" else (1, R - 1)
        err = np.abs(got[lo:hi] - want[lo:hi]).max() / np.abs(want).max()
        assert err < 2e-3, (kind, err)
        # exact parts: the copied rows past ntop - 3 are row ntop - 3
        assert np.abs(got[ntop - 3:hi] - got[ntop - 4]).max() == 0.0, kind


# ------------------------------ rhsp composition: a manufactured solution
# Each velocity component is a mean mode and one spanwise mode,
#   f(x, y, z) = F0(x, y) + F1(x, y) exp(i beta z) + c.c.,
# with F0, F1 sums of separable terms c X(x) Y(y) whose derivatives are known
# in closed form; products go to spanwise modes 0, 1, 2 (no aliasing on
# nz = 24).  A factor is (value, first, second derivative).
def _sin(a, ph):
    return (lambda t: np.sin(a * t + ph), lambda t: a * np.cos(a * t + ph), lambda t: -a * a * np.sin(a * t + ph))


def _one():
    return (np.ones_like, np.zeros_like, np.zeros_like)


def _tanh(c):
    return (lambda t: np.tanh(c * t), lambda t: c * (1 - np.tanh(c * t) ** 2),
            lambda t: -2 * c * c * np.tanh(c * t) * (1 - np.tanh(c * t) ** 2))


def _wall(c):
    # t exp(-c t): zero at the wall, as interpyy's no-slip row 1 assumes
    return (lambda t: t * np.exp(-c * t), lambda t: np.exp(-c * t) * (1 - c * t),
            lambda t: np.exp(-c * t) * (c * c * t - 2 * c))


def _fac(f, t, d):
    if isinstance(f[0], tuple):          # product of two factors
        a, b = f
        return [a[0](t) * b[0](t), a[1](t) * b[0](t) + a[0](t) * b[1](t),
                a[2](t) * b[0](t) + 2 * a[1](t) * b[1](t) + a[0](t) * b[2](t)][d]
    return f[d](t)


def _val(M, X, Y, dx=0, dy=0):
    out = np.zeros(np.broadcast(X, Y).shape, complex)
    for c, fx, gy in M:
        out += c * _fac(fx, X, dx) * _fac(gy, Y, dy)
    return out


def _zprod(F, G):
    """Spanwise modes 0, 1, 2 of the product of F = [F0, F1] and G."""
    def mul(A, B):
        return [(ca * cb, (fa, fb), (ga, gb)) for ca, fa, ga in A for cb, fb, gb in B]

    def conj(A):
        return [(np.conj(c), f, g) for c, f, g in A]
    return [mul(F[0], G[0]) + mul(F[1], conj(G[1])) + mul(conj(F[1]), G[1]),
            mul(F[0], G[1]) + mul(F[1], G[0]), mul(F[1], G[1])]


def _mms_errors(oracle_lib, nx, ny):
    """Largest error of the oracle's right-hand side (rhsp.F90:120-418, the
    rhs fields after rhsp) against d(u_j u)/dx_j - nu (d2/dx2 + d2/dz2) u of
    the manufactured field, each relative to the largest exact term: planes
    2..nx-1 (plane 1 is the inflow, the outflow plane takes the convective
    outflow condition), rows 2..ny of u, w and 2..ny-1 of v (the rows the
    boundary conditions do not overwrite)."""
    cfg = config.small(nx=nx, ny=ny, nz=24, kzmax=0)
    st = dict(ic.make_state(cfg, with_perturbation=False))
    y = np.asarray(st["y"])
    h = cfg.ax * math.pi / (cfg.nx + 1)
    x = np.arange(cfg.nx + 2) * h
    xu, xv = x[1:cfg.nx + 1], x[1:cfg.nx + 1] - h / 2       # u at x(i); v, w at x(i - 1/2)
    yu, yv = 0.5 * (y[0:ny + 1] + y[1:ny + 2]), y[1:ny + 1]  # u, w rows; v rows
    a = 2 * math.pi / x[-1]
    U = [[(1.0, _sin(0.5 * a, 0.3), _tanh(2.0)), (0.2, _one(), _wall(0.5))],
         [(0.1 + 0.05j, _sin(a, 1.0), _wall(0.7))]]
    V = [[(0.05, _sin(a, 0.2), _wall(0.6))], [(0.03 - 0.1j, _sin(0.5 * a, 0.5), _wall(0.8))]]
    W = [[(0.04, _sin(0.7 * a, 0.1), _wall(0.9))], [(0.06 + 0.02j, _sin(a, 2.0), _wall(0.5))]]
    # the reference's spanwise wavenumber is k/az rounded to FP32 (bl_1.F90:360)
    beta = ic.kaz_values(cfg)[1]
    nu = 1.0 / st["re"]

    def put(F, xs, ys):
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        out = np.zeros((cfg.nx, len(ys), cfg.mz), complex)
        out[:, :, 0] = _val(F[0], X, Y).real
        out[:, :, 1] = _val(F[1], X, Y)
        return out

    st.update(u=put(U, xu, yu), v=put(V, xv, yv), w=put(W, xv, yu))
    st["p"] = np.zeros_like(st["v"])
    ob = oracle_lib.OracleBox(cfg, st)
    ob.rhsp(1)

    def exact(cx, cy, cz, F, xs, ys):
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        out = np.zeros((cfg.nx, len(ys), 3), complex)
        for k in range(3):
            t = _val(cx[k], X, Y, dx=1) + _val(cy[k], X, Y, dy=1) + 1j * k * beta * _val(cz[k], X, Y)
            if k < 2:
                t -= nu * (_val(F[k], X, Y, dx=2) - (k * beta) ** 2 * _val(F[k], X, Y))
            out[:, :, k] = t
        return out

    UU, UV, UW, VV, VW, WW = (_zprod(*p) for p in ((U, U), (U, V), (U, W), (V, V), (V, W), (W, W)))
    want = {"rhsupa": exact(UU, UV, UW, U, xu, yu), "rhsvpa": exact(UV, VV, VW, V, xv, yv),
            "rhswpa": exact(UW, VW, WW, W, xv, yu)}
    errs = {}
    for n, wv in want.items():
        rows = slice(1, ny - 1) if n == "rhsvpa" else slice(1, ny)
        g = ob.get(n)[1:-1, rows, :3]
        errs[n] = float(np.abs(g - wv[1:-1, rows]).max() / np.abs(wv[1:-1, rows]).max())
    return errs


def test_rhs_manufactured_solution_converges(oracle_lib):
    """The composition of rhsp: on a manufactured field, the oracle's
    right-hand side of all three components converges to the analytic
    d(u_j u_i)/dx_j - nu (d2/dx2 + d2/dz2) u_i under grid refinement (x and y
    spacing halved twice, the tanh grid nested), at second order or better
    (the 3-point x outflow closures and the stretched-grid y schemes are
    second order), boundary-closure rows included.  A wrong term, sign,
    staggering or buffer would leave an error that does not shrink."""
    levels = [(32, 40), (65, 79), (131, 157)]
    errs = [_mms_errors(oracle_lib, nx, ny) for nx, ny in levels]
    for n in errs[0]:
        e = [lv[n] for lv in errs]
        assert e[0] < 2e-3, (n, e)
        assert e[0] / e[1] > 3.2 and e[1] / e[2] > 3.2, (n, e)


def test_rk3_step_amplification(oracle_lib):
    """The three rhsp substeps of one RK3 step (bl_1.F90:94-125) on one
    spanwise mode that is constant in x and y: only the explicit z-viscous
    term acts at the box centre, so u_m = u_{m-1} + dt (gamma_m L u_{m-1} +
    zeta_m L u_{m-2}) with gamma = (8/15, 5/12, 3/4), zeta = (0, -17/60,
    -5/12) (coeft.f90:494-508) must multiply the mode by the third-order
    Taylor polynomial 1 + z + z^2/2 + z^3/6 of z = -nu kz^2 dt.  This pins
    the RK constants and the rhs / previous-rhs bookkeeping across substeps
    (rhsp.F90:448-475).  The pressure projection between substeps is left
    out (its own pins cover it): the no-slip ghost rows and the recycled
    inflow make this field's divergence non-zero near the walls and the
    inflow, and its projection would reach the centre."""
    cfg = config.small(nx=64, kzmax=0)
    st = dict(ic.make_state(cfg, with_perturbation=False))
    k = cfg.nz2
    for n in ("u", "v", "w"):
        st[n] = np.zeros_like(st[n])
    eps = 1e-6
    st["u"][:, :, k] = eps
    ob = oracle_lib.OracleBox(cfg, st, freeze_dt=True)
    for m in (1, 2, 3):
        ob.rhsp(m)
        ob.outflow_correction()
    dt = ob.time()[1]
    z = -(ic.kaz_values(cfg)[k] ** 2) * dt / st["re"]
    g = 1 + z + z * z / 2 + z ** 3 / 6
    assert abs(z) > 0.05          # the z^2 and z^3 terms are well above the tolerance
    u = ob.get("u")
    i, j = cfg.nx // 2, cfg.ny // 2
    assert u[i, j, k].real == pytest.approx(g * eps, rel=1e-12)
    assert abs(u[i, j, k].real - math.exp(z) * eps) > 1e-5 * eps
