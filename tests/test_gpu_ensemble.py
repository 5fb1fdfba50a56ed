"""Ensembles of independent boxes advanced together (otbl_ensemble_*): each
RK3 substep of all the boxes is one captured graph whose branches are the
boxes' substeps.

* The ensemble gives each box bit for bit what the box gives alone (the same
  kernels in the same order on the same data), in FP64 and in the mixed
  precision, with boxes of different sizes and z transforms (the register
  engine and the LDS Stockham kernels).
* Boxes stepped alone between ensemble steps (buffer roles out of step with
  each other) take a graph of their own role assignment.
* The uneven eight-box ensemble (nx 256..1024, ny 96..192, nz 64..256,
  Re_theta 300..650, seeds 10..17, perturbation 1e-2, FP32 bulk): one RK3
  step of each box against the CPU FP64 oracle.
"""
import time

import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import Ensemble, GpuBox, OTBL_MIXED
from parity_util import rel

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def fields(box):
    return {n: box.get(n) for n in ("u", "v", "w", "p")}


def assert_identical(a, b):
    fa, fb = fields(a), fields(b)
    for n in fa:
        assert np.array_equal(fa[n], fb[n]), (n, rel(fa[n], fb[n]))
    assert a.time() == b.time()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_matches_boxes_alone(precision):
    prec = OTBL_MIXED if precision == "mixed" else 0
    if prec == OTBL_MIXED:   # nx - 1 >= 129 for the mixed Poisson x transforms
        cfgs = [config.small(name="e0", nx=130, ny=40, nz=96, Lx=8.0 * 130 / 32, kzmax=8),
                config.small(name="e1", nx=160, ny=56, nz=256, Lx=8.0 * 160 / 32, kzmax=20, seed=3),
                config.small(name="e2", nx=200, ny=48, nz=160, Lx=8.0 * 200 / 32, kzmax=12, seed=4)]
    else:
        cfgs = [config.small(name="e0"), config.small(name="e1", nx=48, ny=56, nz=36, kzmax=6, seed=3),
                config.small(name="e2", nx=130, ny=40, nz=256, Lx=8.0 * 130 / 32, kzmax=20, seed=4)]
    states = [ic.make_state(c) for c in cfgs]
    ens_boxes = [GpuBox(c, s, precision=prec, graphs=True) for c, s in zip(cfgs, states)]
    alone = [GpuBox(c, s, precision=prec, graphs=True) for c, s in zip(cfgs, states)]
    ens = Ensemble(ens_boxes)
    ens.step(2)
    ens.sync()
    for b in alone:
        b.step(2)
    for a, b in zip(ens_boxes, alone):
        assert_identical(a, b)
    ens.close()
    for b in ens_boxes + alone:
        b.close()


def test_ensemble_with_boxes_out_of_step():
    """Boxes stepped alone between two ensemble steps: box 0 one whole step
    (three substeps), box 1 two substeps.  The buffer roles flip at every
    substep, so the two boxes' roles now differ, and the next ensemble step
    captures a graph for that role assignment."""
    cfgs = [config.small(name="o0"), config.small(name="o1", nx=40, ny=48, nz=30, kzmax=5, seed=5)]
    states = [ic.make_state(c) for c in cfgs]
    eb = [GpuBox(c, s, graphs=True) for c, s in zip(cfgs, states)]
    ref = [GpuBox(c, s, graphs=True) for c, s in zip(cfgs, states)]
    ens = Ensemble(eb)
    ens.step(1)
    eb[0].step(1)        # box 0 alone: roles flipped three times
    eb[1].substep(1)     # box 1 alone: flipped twice (now out of step with box 0)
    eb[1].substep(2)
    ens.sync()
    ens.step(1)
    ens.sync()
    # the same call sequences on boxes stepped alone
    ref[0].step(3)
    ref[1].step(1)
    ref[1].substep(1)
    ref[1].substep(2)
    ref[1].step(1)
    for a, b in zip(eb, ref):
        fa, fb = fields(a), fields(b)
        for n in fa:
            assert np.array_equal(fa[n], fb[n]), n
    ens.close()
    for b in eb + ref:
        b.close()


@pytest.mark.timeout(900)
def test_zpg_uneven_ensemble_vs_oracle(oracle_lib):
    """The eight-box ensemble in the mixed precision, one RK3 step, each box
    against the oracle: the velocity within 1e-6 of the free stream (ten
    times inside the mixed substep's 1e-5), the pressure printed."""
    cfgs = config.ensemble_configs()
    t0 = time.perf_counter()
    states = [ic.make_state(c) for c in cfgs]
    log(f"ensemble states in {time.perf_counter() - t0:.1f} s")
    boxes = [GpuBox(c, s, precision=OTBL_MIXED, graphs=True) for c, s in zip(cfgs, states)]
    log("ensemble paths", [b.paths() for b in boxes])
    ens = Ensemble(boxes)
    ens.step(1)
    ens.sync()
    worst = 0.0
    for c, s, b in zip(cfgs, states, boxes):
        t0 = time.perf_counter()
        ob = oracle_lib.OracleBox(c, s)
        ob.step()
        fo, fg = fields(ob), fields(b)
        scale = max(np.abs(fo[n]).max() for n in ("u", "v", "w"))
        vel = max(float(np.abs(fg[n] - fo[n]).max()) for n in ("u", "v", "w")) / scale
        log(f"{c.name} {c.nx}x{c.ny}x{c.nz}: vel {vel:.2e}, p {rel(fg['p'], fo['p']):.2e} "
            f"(oracle {time.perf_counter() - t0:.1f} s)")
        assert np.isfinite(vel)
        worst = max(worst, vel)
        del ob
    ens.close()
    for b in boxes:
        b.close()
    assert worst < 1e-6, worst
