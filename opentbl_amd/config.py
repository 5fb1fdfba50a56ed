"""Box configurations for the zero-pressure-gradient boundary-layer RK3 step.

Lengths are in units of the inlet boundary-layer thickness d0 (delta_99 of the
inlet profile = 1) and velocities in units of U_inf.  The reference reads its
physical parameters from hrem_1.dat (bl_1.F90:203-244: Re, ax, ay, az, cfl)
with the x length ax*pi and the z length 2*pi*az (escru.F90:133,
bl_1.F90:360); here they are derived from the box sizes:

    ax = Lx / pi,   az = Lz / (2 pi)

Re is chosen so that the reference's inlet_retheta (bl_1.F90:372-396) on the
generated inlet profile returns the requested inlet Re_theta.
"""
from dataclasses import dataclass, field, replace
import math


@dataclass(frozen=True)
class BoxConfig:
    name: str
    nx: int
    ny: int
    nz: int
    Lx: float
    Ly: float
    Lz: float
    re_theta: float
    amp: float            # perturbation amplitude (U_inf units, peak z-rms)
    kzmax: int            # highest perturbed spanwise mode index
    spectrum: str         # "gauss" (flat, Gaussian envelope) or "k53"
    seed: int
    steps: int
    cfl: float = 0.5
    stretch: float = 2.5  # tanh stretching factor of the y grid
    xout_frac: float = 0.4  # recycling plane position as a fraction of nx

    @property
    def nz1(self) -> int:
        return 2 * (self.nz // 3)

    @property
    def nz2(self) -> int:
        return self.nz1 // 2 - 1

    @property
    def mz(self) -> int:
        return self.nz2 + 1

    @property
    def ax(self) -> float:
        return self.Lx / math.pi

    @property
    def az(self) -> float:
        return self.Lz / (2.0 * math.pi)

    @property
    def xout(self) -> int:
        return max(5, min(self.nx - 5, int(round(self.xout_frac * self.nx))))

    @property
    def points(self) -> int:
        """Grid points counted by the throughput metric: nx*ny*nz."""
        return self.nx * self.ny * self.nz

    def check(self):
        """Validate sizes the way the reference's compile-time layout needs
        them (mod.F90:22-27) and the way the GPU kernels need them."""
        if self.nx < 16 or self.ny < 24 or self.nz < 12:
            raise ValueError(f"{self.name}: box too small ({self.nx}x{self.ny}x{self.nz})")
        if self.nz % 2:
            raise ValueError(f"{self.name}: nz must be even")
        if not (1 < self.xout < self.nx):
            raise ValueError(f"{self.name}: xout={self.xout} outside the grid")
        if self.kzmax > self.nz2:
            raise ValueError(f"{self.name}: kzmax={self.kzmax} above nz2={self.nz2}")
        return self


CONFIGS = {
    # plain FP64 reference method, CPU-runnable
    "zpg256x96x64": BoxConfig("zpg256x96x64", 256, 96, 64, 60.0, 4.0, 8.0, 300.0, 1e-3, 8, "gauss", 0, 20),
    # FP64 GPU drop-in, production-like single-GPU size
    "zpg1024x192x512": BoxConfig("zpg1024x192x512", 1024, 192, 512, 60.0, 4.0, 8.0, 600.0, 0.05, 32, "k53", 1, 50),
    # mixed precision headline box
    "zpg1536x315x1024": BoxConfig("zpg1536x315x1024", 1536, 315, 1024, 90.0, 4.0, 12.0, 1000.0, 0.1, 64, "k53", 2, 100),
    # largest single-GPU box
    "zpg3072x360x2048": BoxConfig("zpg3072x360x2048", 3072, 360, 2048, 180.0, 4.0, 24.0, 1500.0, 0.1, 128, "k53", 3, 200),
}


def ensemble_configs():
    """The uneven eight-box ensemble: nx, ny, nz and Re_theta per box,
    seeds 10-17, perturbation 1e-2."""
    nxs = [256, 384, 512, 640, 768, 896, 1024, 1024]
    nys = [96, 112, 128, 144, 160, 176, 192, 192]
    nzs = [64, 96, 128, 160, 192, 224, 256, 256]
    rts = [300, 350, 400, 450, 500, 550, 600, 650]
    out = []
    for b in range(8):
        out.append(BoxConfig(f"c4_{b}", nxs[b], nys[b], nzs[b], 60.0, 4.0, 8.0,
                             float(rts[b]), 1e-2, min(8, 2 * (nzs[b] // 3) // 2 - 1),
                             "gauss", 10 + b, 20))
    return out


def small(name="tiny", nx=32, ny=40, nz=24, **kw):
    """A small box for parity tests (the oracle finishes it in well under a
    second per substep)."""
    base = BoxConfig(name, nx, ny, nz, 8.0, 4.0, 2.0, 300.0, 0.05, 4, "gauss", 7, 3)
    return replace(base, **kw).check()
