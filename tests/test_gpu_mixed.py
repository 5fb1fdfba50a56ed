"""Mixed-precision substep (OTBL_MIXED) against the CPU FP64 oracle and the
FP64 GPU path.

In the mixed substep the fields, the compact derivatives, the z transforms,
the products and the Poisson x transforms are FP32; the energies, the CFL
maxima, the recycling statistics, the mass-flux correction, the mean-flow
(kz = 0) pressure lines and the residual of one refinement step of the other
modes' y Poisson solves are FP64; so is the state of the mean-flow u and
p.  Bar: the velocity (u, v, w) within 1e-6 of the FP64 oracle after whole
RK3 steps, relative to the largest velocity component (the free stream), ten
times inside the 1e-5 the mixed substep is specified to; measured: ~3e-8
after four steps.  Each component is also printed relative to its
own largest magnitude: v and w carry the FP32 rounding of u (~6e-8 of U_inf,
spread by the projection) and so stand higher against their own, smaller,
maxima.  The pressure solve itself is held to 1e-5 of max |p| on identical
input; after whole steps p is only printed: it is (1/dt) times the potential
of the velocity's divergence, so it carries the FP32 velocity differences
divided by dt.
"""
import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import GpuBox, OTBL_MIXED
from parity_util import rel

pytestmark = pytest.mark.gpu

VEL_TOL = 1e-6
P_TOL = 1e-5


# nz = 1024 (radix 16-16-4) and 2048 (a two-wave team) are the z lengths of
# the 1536 x 315 x 1024 and 3072 x 360 x 2048 boxes; nz = 64, 96, 160, 224
# (factors 2, 3, 5, 7) run on the LDS Stockham transforms, the z lengths of
# the ensemble's small boxes
@pytest.fixture(scope="module", params=[(130, 256), (257, 512), (130, 1024), (130, 2048), (130, 64), (130, 96),
                                        (130, 160), (130, 224)],
                ids=["x130z256", "x257z512", "x130z1024", "x130z2048", "x130z64", "x130z96", "x130z160",
                     "x130z224"])
def mbox(request):
    nx, nz = request.param
    cfg = config.small(name=f"mix{nx}", nx=nx, ny=40, nz=nz, Lx=8.0 * nx / 32, kzmax=min(40, nz // 3 - 1))
    return cfg, ic.make_state(cfg)


def errors(a, b):
    """{"vel": max |d(u, v, w)| / max |(u, v, w)| of b, "p": ..., and each
    component relative to its own maximum}."""
    fa = {n: a.get(n) for n in ("u", "v", "w", "p")}
    fb = {n: b.get(n) for n in ("u", "v", "w", "p")}
    scale = max(np.abs(fb[n]).max() for n in ("u", "v", "w"))
    errs = {"vel": max(float(np.abs(fa[n] - fb[n]).max()) for n in ("u", "v", "w")) / scale,
            "p": rel(fa["p"], fb["p"])}
    errs.update({n + "_own": rel(fa[n], fb[n]) for n in ("u", "v", "w")})
    return errs


def check(errs, with_p=False):
    assert errs["vel"] < VEL_TOL, errs
    if with_p:
        assert errs["p"] < P_TOL, errs


def test_mixed_paths(mbox):
    cfg, st = mbox
    gm = GpuBox(cfg, st, precision=OTBL_MIXED)
    paths = gm.paths()
    gm.close()
    zpath = "engine" if cfg.nz >= 256 else "stockham"
    assert paths["y"].startswith("chunk") and paths["z"].startswith(zpath), paths
    assert paths["dct"].startswith("engine") and paths["x"] == "window", paths


def test_mixed_fields_round_to_fp32(mbox):
    """Fields are held in FP32 except the mean-flow (mode 0, real part) u and
    p, which are held in FP64."""
    cfg, st = mbox
    gm = GpuBox(cfg, st, precision=OTBL_MIXED)
    got = {n: gm.get(n) for n in ("u", "v", "w", "p")}
    gm.close()
    for n, g in got.items():
        a = np.asarray(st[n])
        want = a.astype(np.complex64).astype(np.complex128)
        if n in ("u", "p"):
            want[..., 0] = a[..., 0].real + 1j * want[..., 0].imag
        assert np.array_equal(g, want), n


def test_mixed_step_vs_oracle(mbox, oracle_lib):
    cfg, st = mbox
    ob = oracle_lib.OracleBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED)
    ob.step()
    gm.step(1)
    errs = errors(gm, ob)
    eo, eg = ob.ener(), gm.ener()
    gm.close()
    print("mixed vs oracle after 1 step", errs, flush=True)
    check(errs)
    assert rel(eg, eo) < 1e-6


def test_mixed_vs_fp64_steps(mbox):
    cfg, st = mbox
    gd = GpuBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED)
    gd.step(3)
    gm.step(3)
    errs = errors(gm, gd)
    td, tm = gd.time(), gm.time()
    gd.close()
    gm.close()
    print("mixed vs FP64 after 3 steps", errs, flush=True)
    check(errs)
    assert abs(tm[0] - td[0]) <= 1e-5 * abs(td[0])


def test_mixed_graphs_match_eager(mbox):
    cfg, st = mbox
    ge = GpuBox(cfg, st, precision=OTBL_MIXED)
    gg = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    ge.step(2)
    gg.step(2)
    errs = {n: rel(gg.get(n), ge.get(n)) for n in ("u", "v", "w", "p")}
    ge.close()
    gg.close()
    assert all(e == 0.0 for e in errs.values()), errs


def test_mixed_pressure_solve(mbox):
    """One pressure projection of a divergent field: the mixed solve (FP32
    modes with the FP64-residual refinement, FP64 mean-flow lines) against
    the FP64 solve."""
    cfg, st = mbox
    rng = np.random.default_rng(7)
    gd = GpuBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED)
    for n in ("u", "v", "w"):
        a = np.asarray(st[n])
        f = a + 1e-2 * np.abs(a).max() * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape))
        gd.set(n, f)
        gm.set(n, f)
    gd.pois(0.5)
    gm.pois(0.5)
    errs = errors(gm, gd)
    gd.close()
    gm.close()
    print("mixed vs FP64 projection", errs, flush=True)
    check(errs, with_p=True)


def test_mixed_rejects_unsupported():
    """A box without the power-of-two Bluestein length (nx - 1 < 129) has no
    mixed substep; creation fails with a message."""
    cfg = config.small()
    st = ic.make_state(cfg)
    with pytest.raises(Exception, match="OTBL_MIXED"):
        GpuBox(cfg, st, precision=OTBL_MIXED)


def test_mixed_long_x(oracle_lib):
    """A box whose Poisson x transforms take the 3 x 1024 Bluestein length
    (nx - 1 = 1099) and whose x lines are split into segments: one mixed step
    against the oracle, two against the FP64 steps on the GPU."""
    cfg = config.small(name="mixlong", nx=1100, ny=40, nz=256, Lx=8.0 * 1100 / 32, kzmax=40)
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    gd = GpuBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    pm = gm.paths()
    assert pm["dct"] == "engine3072" and int(pm["xseg"]) > 1, pm
    ob.step()
    gm.step(1)
    e_o = errors(gm, ob)
    gd.step(2)
    gm.step(1)
    errs = errors(gm, gd)
    gd.close()
    gm.close()
    print("mixed vs oracle, nx=1100, after 1 step", e_o, "vs FP64 after 2 steps", errs, flush=True)
    check(e_o)
    check(errs)


@pytest.mark.parametrize("ny,chunk", [(150, 12), (315, 20)], ids=["y150", "y315"])
def test_mixed_tall_columns_vs_fp64(ny, chunk):
    """Columns of the production height (ny = 315) and of an intermediate
    one: the y lines take 20 (12) rows per chunk and the FP32 rk_final its
    half-wave chunking of 10 (6) rows; mixed steps against the FP64 steps on
    the GPU."""
    cfg = config.small(name=f"mixy{ny}", nx=130, ny=ny, nz=256, Lx=8.0 * 130 / 32, kzmax=40)
    st = ic.make_state(cfg)
    gd = GpuBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    pm = gm.paths()
    assert pm["y"] == f"chunk{chunk}h", pm
    gd.step(2)
    gm.step(2)
    errs = errors(gm, gd)
    gd.close()
    gm.close()
    print(f"mixed vs FP64, ny={ny}, after 2 steps", errs, flush=True)
    check(errs)


def test_mixed_full_wave_rk_final(mbox, monkeypatch):
    """OTBL_Y_HALF=0 keeps the full-wave FP32 rk_final (64 columns per
    workgroup, 16 chunks): the same steps as the half-wave kernel up to FP32
    rounding (other chunk boundaries), both within the bar of the FP64
    path."""
    cfg, st = mbox
    gh = GpuBox(cfg, st, precision=OTBL_MIXED)
    monkeypatch.setenv("OTBL_Y_HALF", "0")
    gf = GpuBox(cfg, st, precision=OTBL_MIXED)
    monkeypatch.delenv("OTBL_Y_HALF")
    gd = GpuBox(cfg, st)
    ph, pf = gh.paths(), gf.paths()
    assert ph["y"].endswith("h") and not pf["y"].endswith("h"), (ph, pf)
    for g in (gh, gf, gd):
        g.step(2)
    e_hf, e_fd = errors(gh, gf), errors(gf, gd)
    for g in (gh, gf, gd):
        g.close()
    print("half-wave vs full-wave rk_final", e_hf, "full-wave vs FP64", e_fd, flush=True)
    check(e_hf)
    check(e_fd)
