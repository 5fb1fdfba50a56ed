"""Host-side diagnostics of the fields a box returns (complex arrays
(nx, R, nz2+1) as GpuBox.get gives them): the discrete divergence of the
pressure step and the net mass flux through the box.  Used by the tests
and by bench.py to state the divergence of each run."""
import numpy as np


def divergence_rel(cfg, y, u, v, w, chunk=64):
    """max |div| / (max |u| / dx) over planes 2..nx and rows 1..ny-1, with the
    discrete divergence of pois.f90:41-54 (ikz w + dv/dy + du/dx on the
    staggered grid), evaluated in runs of planes to bound host memory."""
    from opentbl_amd import ic
    kaz = ic.kaz_values(cfg)
    dy = np.diff(np.asarray(y))
    idx = (cfg.nx + 1) / (cfg.ax * np.pi)   # 1/dx(1), coeft.f90:397-424
    ny = cfg.ny
    worst = 0.0
    for i0 in range(1, cfg.nx, chunk):
        i1 = min(cfg.nx, i0 + chunk)
        d = (w[i0:i1, 1:ny, :] * (1j * kaz[None, None, :])
             + (v[i0:i1, 1:ny, :] - v[i0:i1, 0:ny - 1, :]) / dy[1:ny][None, :, None]
             + idx * (u[i0:i1, 1:ny, :] - u[i0 - 1:i1 - 1, 1:ny, :]))
        worst = max(worst, float(np.abs(d).max()))
    return worst / (np.abs(u).max() * idx)


def net_flux_rel(cfg, y, u, v):
    """The net mass flux through the box of boun.f90:63-93 (top v times dx,
    outlet minus inlet u times dy, mode 0) relative to the inflow flux."""
    dy = np.diff(np.asarray(y))
    ny = cfg.ny
    h0 = cfg.ax * np.pi / (cfg.nx + 1)
    top = float(v[1:, ny - 1, 0].real.sum()) * h0
    inl = float((u[0, 1:ny, 0].real * dy[1:ny]).sum())
    out = float((u[-1, 1:ny, 0].real * dy[1:ny]).sum())
    return abs(top - inl + out) / abs(inl)
