"""Grid and initial-condition generator for the boundary-layer box.

The reference has no generator on its main path: getstartzy reads the y grid,
u, v, w, p, um and the times from restart files (getstartzy-bg.F90:180-295),
and sets the inflow profiles u0, v0 from plane 1 (getstartzy-bg.F90:274-275).
This module writes that state from scratch:

* y grid: tanh stretching clustered at the wall, y(1) = 0 at the wall and
  y(ny) = Ly at the top, with mirror ghost rows y(0) = -y(2) and
  y(ny+1) = 2 y(ny) - y(ny-1) (the reference's y(0:ny+1), mod.F90:145).
* mean flow: a Blasius layer whose delta_99 grows from 1 at the inlet as
  delta^2 = 1 + eta99^2 x / Re; v is integrated from the staggered
  continuity equation (the divergence of pois.f90:41-54), so the mean flow is
  discretely divergence-free at planes 2..nx.
* perturbation: spanwise modes 1..kzmax built from two streamfunctions (one in
  the y-z plane for v, w and one in the x-z plane for u, w), each
  discretely divergence-free under the same staggered operator, with a flat
  (Gaussian envelope) or kz^-5/3 energy spectrum, seeded.

Layout of the returned arrays (the planes layout of the reference): complex
fields have numpy shape (nx, R, nz2+1) = (plane i, row j, mode k), so the
real view (nx, R, nz1) is the reference's complex*16 (0:nz2, R, nx) memory.
"""
import math
import numpy as np

from .config import BoxConfig


def ygrid(cfg: BoxConfig) -> np.ndarray:
    """y(0:ny+1) as a float64 array of length ny+2."""
    ny = cfg.ny
    eta = np.arange(ny, dtype=np.float64) / (ny - 1)
    g = cfg.stretch
    if g > 0:
        yv = cfg.Ly * (1.0 + np.tanh(g * (eta - 1.0)) / math.tanh(g))
    else:                      # stretch 0: uniform grid
        yv = cfg.Ly * eta
    yv[0] = 0.0
    y = np.empty(ny + 2)
    y[1:ny + 1] = yv
    y[0] = -y[2]
    y[ny + 1] = 2.0 * y[ny] - y[ny - 1]
    return y


_BLASIUS = None


def blasius():
    """Blasius solution f(eta), f'(eta) on a fine eta grid (RK4 shooting on
    f''' + f f''/2 = 0, f(0) = f'(0) = 0, f'(inf) = 1)."""
    global _BLASIUS
    if _BLASIUS is not None:
        return _BLASIUS
    h, n = 1e-3, 15000

    def integrate(fpp0):
        s = np.array([0.0, 0.0, fpp0])
        out = np.empty((n + 1, 3))
        out[0] = s

        def rhs(v):
            return np.array([v[1], v[2], -0.5 * v[0] * v[2]])
        for i in range(n):
            k1 = rhs(s)
            k2 = rhs(s + 0.5 * h * k1)
            k3 = rhs(s + 0.5 * h * k2)
            k4 = rhs(s + h * k3)
            s = s + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
            out[i + 1] = s
        return out

    # f''(0) of the Blasius problem (Howarth's value, 0.332057336215196)
    sol = integrate(0.332057336215196)
    eta = np.arange(n + 1) * h
    eta99 = float(np.interp(0.99, sol[:, 1], eta))
    _BLASIUS = (eta, sol[:, 0], sol[:, 1], eta99)
    return _BLASIUS


def _profile(yq, delta):
    """Blasius U and (eta f' - f) at heights yq for thickness delta."""
    eta, f, fp, eta99 = blasius()
    e = np.abs(yq) * eta99 / delta
    U = np.interp(e, eta, fp, right=1.0)
    F = np.interp(e, eta, eta * fp - f, right=eta[-1] * 1.0 - f[-1])
    return np.sign(yq) * U, F


def inlet_retheta_sum(y, u0, ny):
    """The reference's inlet_retheta sum without the Re factor
    (bl_1.F90:381-388; 0.99 is a single-precision literal there)."""
    dy = np.diff(y)
    uinf = u0[ny + 1 - 20 - 1:ny + 1 - 5].sum() / 16.0
    u99 = float(np.float32(0.99)) * uinf
    jtop = 0
    for j in range(1, ny - 1):
        if u0[j - 1] >= u99:
            break
        jtop = j + 1
    s = 0.0
    for j in range(2, jtop + 4):
        s += (uinf - u0[j - 1]) * u0[j - 1] * dy[j - 1]
    return s / uinf


def kaz_values(cfg: BoxConfig):
    """Spanwise wavenumbers as the reference holds them: cmplx(0d0,k/az) is
    a default (single) precision complex (bl_1.F90:360)."""
    k = np.arange(cfg.nz2 + 1, dtype=np.float64)
    return (k / cfg.az).astype(np.float32).astype(np.float64)


def make_state(cfg: BoxConfig, with_perturbation=True):
    """Return a dict with the full box state:

    y (ny+2), re, u, w (nx, ny+1, Mz) complex, v, p (nx, ny, Mz) complex,
    u0 (ny+1), v0 (ny), um (ny+1).
    """
    cfg.check()
    nx, ny, mz = cfg.nx, cfg.ny, cfg.mz
    y = ygrid(cfg)
    dy = np.diff(y)                          # dy(0:ny)
    dxg = cfg.ax * math.pi / (nx + 1)        # coeft.f90:181,401
    xs = np.arange(nx) * dxg                 # plane positions (x(i) - x(1))
    yu = 0.5 * (y[:-1] + y[1:])              # u rows 1..ny+1
    yv = y[1:ny + 1]                         # v rows 1..ny

    # ---- mean flow: Re from the inlet profile, growing Blasius layer ----
    U_in, _ = _profile(yu, 1.0)
    re = cfg.re_theta / inlet_retheta_sum(y, U_in, ny)
    eta99 = blasius()[3]
    delta = np.sqrt(1.0 + eta99 ** 2 * xs / re)
    U = np.empty((nx, ny + 1))
    for i in range(nx):
        U[i], _ = _profile(yu, delta[i])
    V = np.zeros((nx, ny))
    # inflow v from the Blasius similarity solution, v = 0.5 sqrt(1/(Re x_eff)) (eta f' - f)
    xeff0 = re / eta99 ** 2
    _, F = _profile(yv, 1.0)
    V[0] = 0.5 * math.sqrt(1.0 / (re * xeff0)) * F
    V[0, 0] = 0.0
    idx = 1.0 / dxg
    for i in range(1, nx):
        du = (U[i, 1:ny] - U[i - 1, 1:ny]) * idx   # rows 2..ny of u
        V[i, 1:] = -np.cumsum(dy[1:ny] * du)

    u = np.zeros((nx, ny + 1, mz), np.complex128)
    v = np.zeros((nx, ny, mz), np.complex128)
    w = np.zeros((nx, ny + 1, mz), np.complex128)
    p = np.zeros((nx, ny, mz), np.complex128)
    u[:, :, 0] = U
    v[:, :, 0] = V

    if with_perturbation and cfg.kzmax > 0:
        _add_perturbation(cfg, u, v, w, y, xs, dxg)

    u0 = u[0, :, 0].real.copy()
    v0 = v[0, :, 0].real.copy()
    return dict(y=y, re=re, u=u, v=v, w=w, p=p, u0=u0, v0=v0, um=u0.copy(),
                ax=cfg.ax, az=cfg.az, xout=cfg.xout)


def _add_perturbation(cfg, u, v, w, y, xs, dxg):
    rng = np.random.default_rng(cfg.seed)
    nx, ny = cfg.nx, cfg.ny
    K = cfg.kzmax
    kz = kaz_values(cfg)[1:K + 1]
    k = np.arange(1, K + 1, dtype=np.float64)
    if cfg.spectrum == "k53":
        ak = k ** (-5.0 / 6.0)
    else:
        ak = np.ones(K)
    c1 = (rng.standard_normal(K) + 1j * rng.standard_normal(K)) * ak
    c2 = (rng.standard_normal(K) + 1j * rng.standard_normal(K)) * ak
    m1 = rng.integers(1, 7, K)
    m2 = rng.integers(1, 7, K)
    ph1 = rng.uniform(0, 2 * math.pi, K)
    ph2 = rng.uniform(0, 2 * math.pi, K)
    Lx = cfg.Lx
    gx1 = np.cos(2 * math.pi * np.outer(xs, m1) / Lx + ph1)    # (nx, K)
    gx2 = np.cos(2 * math.pi * np.outer(xs, m2) / Lx + ph2)
    if cfg.spectrum == "gauss":
        bump = np.exp(-((xs - 0.25 * Lx) / (0.15 * Lx)) ** 2)[:, None]
        gx1 = gx1 * bump
        gx2 = gx2 * bump
    sig = 0.5
    yv = y[1:ny + 1]
    psi = (yv / sig) ** 2 * np.exp(-(yv / sig) ** 2)            # at v rows, psi(0)=psi'(0)=0
    yu = 0.5 * (y[:-1] + y[1:])
    chi = (yu / sig) ** 2 * np.exp(-(yu / sig) ** 2)
    chi[0] = 0.0
    chi[ny] = 0.0
    dy = np.diff(y)
    # y-z streamfunction: v = i kz psi, w(r) = -(psi_r - psi_{r-1}) / dy(r-1)
    pv = np.zeros((nx, ny, K), np.complex128)
    pv[:] = (1j * kz * c1)[None, None, :] * psi[None, :, None] * gx1[:, None, :]
    pw = np.zeros((nx, ny + 1, K), np.complex128)
    dpsi = (psi[1:] - psi[:-1]) / dy[1:ny]                     # rows 2..ny
    pw[:, 1:ny, :] = -(c1[None, None, :] * dpsi[None, :, None] * gx1[:, None, :])
    # x-z streamfunction: u = i kz chi, w -= (chi_i - chi_{i-1}) / dx
    chix = chi[None, :, None] * gx2[:, None, :] * c2[None, None, :]   # (nx, ny+1, K)
    pu = 1j * kz[None, None, :] * chix
    dchi = np.zeros_like(chix)
    dchi[1:] = (chix[1:] - chix[:-1]) / dxg
    pw -= dchi
    # peak z-rms of each component -> amp
    def zrms(a):
        return float(np.sqrt(2.0 * (np.abs(a) ** 2).sum(axis=2)).max())
    s = cfg.amp / max(zrms(pu), zrms(pv), zrms(pw), 1e-300)
    u[:, :, 1:K + 1] += s * pu
    v[:, :, 1:K + 1] += s * pv
    w[:, :, 1:K + 1] += s * pw


def real_view(a):
    """Complex (nx, R, Mz) field -> float64 (nx, R, nz1) view."""
    return np.ascontiguousarray(a).view(np.float64)
