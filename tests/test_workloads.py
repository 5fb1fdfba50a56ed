"""The benchmark workloads at their full sizes on the GPU.

* 256 x 96 x 64 (FP64, fixed dt, 20 RK3 steps): the whole trajectory against
  the CPU FP64 oracle at 1e-10 (u, v, w, p; the kz = 0 pressure line at 1e-8,
  see parity_util), Re_theta(x) and Cf(x) at 1e-10.
* 1024 x 192 x 512 (FP64, 50 RK3 steps): size-independent properties, as the
  oracle is too slow at this size: finite fields, the discrete divergence of
  pois.f90:41-54 at round-off, the net mass flux of boun.f90:63-93 zero,
  Re_theta at the inflow plane equal to the inlet value the box was set up
  with, a positive Cf(x).
* 1536 x 315 x 1024 (mixed against FP64, 100 RK3 steps each from the same
  state): the velocity within 1e-5 of the FP64 run relative to the free
  stream (the mixed substep's specification), each field's error printed
  against its own maximum, the divergence at FP32 round-off, Re_theta(x) and
  Cf(x) against the FP64 run.
* 3072 x 360 x 2048 (mixed, then FP64, 2 RK3 steps each; the two boxes do not
  fit the GPU together): the same properties and the mixed run against the
  FP64 run.
"""
import gc
import time

import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import GpuBox, OTBL_DT_FROZEN, OTBL_MIXED
from parity_util import assert_state, divergence_rel, net_flux_rel, rel, state_errors

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def fields(box):
    return {n: box.get(n) for n in ("u", "v", "w", "p")}


def field_errors(a, b):
    """Largest |a - b| of each field relative to its own largest |b|, and the
    velocity error relative to the largest velocity component (U_inf)."""
    scale = max(float(np.abs(b[n]).max()) for n in ("u", "v", "w"))
    errs = {n: rel(a[n], b[n]) for n in ("u", "v", "w", "p")}
    errs["vel"] = max(float(np.abs(a[n] - b[n]).max()) for n in ("u", "v", "w")) / scale
    return errs


def test_zpg256x96x64_fp64_vs_oracle(oracle_lib):
    cfg = config.CONFIGS["zpg256x96x64"]
    st = ic.make_state(cfg)
    ob = oracle_lib.OracleBox(cfg, st, freeze_dt=True)
    gb = GpuBox(cfg, st, dt_mode=OTBL_DT_FROZEN, graphs=True)
    t0 = time.perf_counter()
    for s in range(cfg.steps):
        ob.step()
    log(f"oracle: {cfg.steps} steps in {time.perf_counter() - t0:.1f} s")
    gb.step(cfg.steps)
    errs = state_errors(gb, ob)
    rg, cg = gb.retheta_cf()
    ro, co = ob.retheta_cf()
    tg, to = gb.time(), ob.time()
    gb.close()
    log(f"256x96x64 after {cfg.steps} steps", errs, "Re_theta", rel(rg, ro), "Cf", rel(cg, co))
    assert_state(errs)
    assert rel(rg, ro) < 1e-10 and rel(cg, co) < 1e-10
    assert abs(tg[1] - to[1]) <= 1e-14 * to[1] and abs(tg[0] - to[0]) <= 1e-13 * to[0]


def test_zpg1024x192x512_fp64_properties():
    cfg = config.CONFIGS["zpg1024x192x512"]
    st = ic.make_state(cfg)
    gb = GpuBox(cfg, st, graphs=True)
    rthin = gb.genflu_info()["rthin"]
    gb.step(cfg.steps)
    f = fields(gb)
    rth, cf = gb.retheta_cf()
    gb.close()
    assert all(np.isfinite(a).all() for a in f.values())
    div = divergence_rel(cfg, st["y"], f["u"], f["v"], f["w"])
    flux = net_flux_rel(cfg, st["y"], f["u"], f["v"])
    log(f"1024x192x512 after {cfg.steps} steps: divergence {div:.2e}, net flux {flux:.2e}, "
        f"Re_theta(1) {rth[0]:.6f} (inlet {rthin:.6f}), Cf {cf.min():.3e}..{cf.max():.3e}")
    assert div < 1e-11
    assert flux < 1e-12
    assert abs(rth[0] - rthin) <= 1e-10 * rthin
    assert abs(rthin - cfg.re_theta) <= 1e-3 * cfg.re_theta
    assert (cf > 0).all()


def test_zpg1536x315x1024_mixed_vs_fp64():
    cfg = config.CONFIGS["zpg1536x315x1024"]
    st = ic.make_state(cfg)
    gd = GpuBox(cfg, st, graphs=True)
    gd.step(cfg.steps)
    fd = fields(gd)
    rd, cd = gd.retheta_cf()
    td = gd.time()
    gd.close()
    del gd
    gc.collect()
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    gm.step(cfg.steps)
    fm = fields(gm)
    rm, cm = gm.retheta_cf()
    tm = gm.time()
    gm.close()
    errs = field_errors(fm, fd)
    div_m = divergence_rel(cfg, st["y"], fm["u"], fm["v"], fm["w"])
    div_d = divergence_rel(cfg, st["y"], fd["u"], fd["v"], fd["w"])
    log(f"1536x315x1024 mixed vs FP64 after {cfg.steps} steps: {errs}; Re_theta {rel(rm, rd):.2e}, "
        f"Cf {rel(cm, cd):.2e}; divergence FP64 {div_d:.2e}, mixed {div_m:.2e}; t {tm[0]:.6f} vs {td[0]:.6f}")
    assert all(np.isfinite(a).all() for a in fm.values())
    assert errs["vel"] < 1e-5, errs
    # FP64 round-off of the last projection: 7e-12 at 1024x192x512, 6.7e-11
    # here (longer x transforms, 315 stretched rows)
    assert div_d < 1e-10 and div_m < 1e-5
    assert rel(rm, rd) < 1e-5 and rel(cm, cd) < 1e-3
    assert abs(tm[0] - td[0]) <= 1e-5 * td[0]


def test_zpg3072x360x2048_mixed_and_fp64():
    cfg = config.CONFIGS["zpg3072x360x2048"]
    steps = 2
    t0 = time.perf_counter()
    st = ic.make_state(cfg)
    log(f"3072x360x2048 state in {time.perf_counter() - t0:.1f} s")
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    log("3072x360x2048 mixed paths", gm.paths())
    gm.step(steps)
    fm = fields(gm)
    rm, cm = gm.retheta_cf()
    gm.close()
    del gm
    gc.collect()
    assert all(np.isfinite(a).all() for a in fm.values())
    div_m = divergence_rel(cfg, st["y"], fm["u"], fm["v"], fm["w"])
    flux_m = net_flux_rel(cfg, st["y"], fm["u"], fm["v"])
    log(f"3072x360x2048 mixed after {steps} steps: divergence {div_m:.2e}, net flux {flux_m:.2e}")
    gd = GpuBox(cfg, st, graphs=True)
    del st
    gc.collect()
    gd.step(steps)
    fd = {}
    for n in ("u", "v", "w", "p"):
        fd[n] = gd.get(n)
    rd, cd = gd.retheta_cf()
    gd.close()
    errs = field_errors(fm, fd)
    log(f"3072x360x2048 mixed vs FP64 after {steps} steps: {errs}; Re_theta {rel(rm, rd):.2e}, Cf {rel(cm, cd):.2e}")
    assert div_m < 1e-5 and flux_m < 1e-6
    assert errs["vel"] < 1e-5, errs
    assert rel(rm, rd) < 1e-5 and rel(cm, cd) < 1e-3
