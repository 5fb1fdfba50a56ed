"""The autotuner (otbl_autotune): it times the candidate launch shapes of a
box (x segments, planes per thread of the predictor and projection, FP32
rk_final chunks), keeps the fastest, leaves the state unchanged, and the
tuned box still matches the CPU FP64 oracle (FP64) or its untuned twin
(mixed)."""
import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import GpuBox, OTBL_MIXED
from parity_util import assert_state, rel, state_errors

pytestmark = pytest.mark.gpu


def test_autotune_fp64_vs_oracle(oracle_lib):
    cfg = config.small(name="tune64", nx=500, ny=40, nz=36, Lx=8.0 * 500 / 32, kzmax=6)
    st = ic.make_state(cfg)
    gb = GpuBox(cfg, st, graphs=True)
    before = {n: gb.get(n) for n in ("u", "v", "w", "p")}
    rep = gb.autotune(3)
    print("autotune:", rep, gb.paths(), flush=True)
    assert gb.paths()["tuned"] == "1"
    assert "xseg1=" in rep and "xseg2=" in rep and "prun8=" in rep
    for n, a in before.items():               # the state is untouched
        assert np.array_equal(gb.get(n), a), n
    ob = oracle_lib.OracleBox(cfg, st)
    gb.step(1)
    ob.step()
    errs = state_errors(gb, ob)
    gb.close()
    assert_state(errs)


def test_autotune_mixed_vs_untuned():
    cfg = config.small(name="tune32", nx=300, ny=40, nz=256, Lx=8.0 * 300 / 32, kzmax=20)
    st = ic.make_state(cfg)
    a = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    b = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    rep = b.autotune(3)
    print("autotune:", rep, b.paths(), flush=True)
    assert "yhalf1=" in rep and "yhalf0=" in rep
    a.step(2)
    b.step(2)
    scale = max(np.abs(a.get(n)).max() for n in ("u", "v", "w"))
    err = max(np.abs(b.get(n) - a.get(n)).max() for n in ("u", "v", "w")) / scale
    print("tuned vs untuned mixed, 2 steps:", err, "p", rel(b.get("p"), a.get("p")), flush=True)
    a.close()
    b.close()
    assert err < 1e-6
