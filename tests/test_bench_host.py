"""Host logic of bench.py (no GPU): which kernel the roofline is quoted on,
its bound, and the hardware byte counts read from the committed rocprofv3
summaries."""
import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from opentbl_amd import config  # noqa: E402

CFG = config.CONFIGS["zpg1536x315x1024"]


def rows_for(model, times):
    """report()'s rows from per-substep launch counts and total times (ms)."""
    rep = {k: (n, ms) for k, (n, ms) in times.items()}
    return bench.report("test", rep, model)


def test_model_splits_the_substep_fields():
    """rk_final moves 5 fields of one component per launch (res, rhs,
    rhs_prev, u read; u written): over its three launches (u, v, w) that is
    5 fields of (nx, ny+1, nz1) elements twice and of (nx, ny, nz1) once."""
    m = bench.kernel_model(CFG, 4)
    by, fl, n = m["rk_final"]
    X, z = CFG.nx, CFG.nz1
    field_u, field_v = X * (CFG.ny + 1) * z * 4, X * CFG.ny * z * 4
    assert n == 3 and fl == 0.0
    assert by * 3 == pytest.approx(5 * (2 * field_u + field_v))
    # one complex transform of nz points per pair of real rows: products1
    # runs 5 per row j <= ny and 3 for row ny + 1
    _, fl1, _ = m["products1"]
    per = 5 * CFG.nz * math.log2(CFG.nz)
    assert fl1 == pytest.approx(X * (5 * CFG.ny + 3) * per)


def test_roofline_takes_the_dominant_kernel_and_its_bound():
    m = bench.kernel_model(CFG, 4)
    rows = rows_for(m, {"products1": (1, 4.4), "rk_final": (3, 4.3), "interpyy": (3, 1.8)})
    r = bench.roofline(rows, m, CFG.name, "mixed")
    assert r["kernel"] == "products1"
    # 11.7 flops per byte: under the 157.3 TFLOP/s : 8 TB/s ridge, so the HBM roof
    by, fl, _ = m["products1"]
    assert fl / by < bench.FP32_PEAK_TFLOPS / (bench.HBM_PEAK_GBS / 1e3)
    assert r["bound"] == "hbm" and r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS)
    h = bench.roofline(rows, m, CFG.name, "mixed", hbm_only=True)
    assert h["kernel"] == "rk_final"
    assert h["achieved"] == pytest.approx(m["rk_final"][0] / (4.3 / 3 * 1e-3) / 1e9)


@pytest.mark.parametrize("kname,prec", [("rk_final", "mixed"), ("rk_tail", "fp64"), ("project", "mixed"),
                                        ("save_predict", "mixed"), ("products1", "mixed"), ("products1", "fp64")])
def test_profiled_traffic_finds_the_kernel(kname, prec):
    """The committed summaries name the HIP functions (rk_final_yh<10> in
    FP32, the fused tail rk_tail_y<20, double, k> in FP64, the pair kernels
    of the FP32 element-wise steps); each lookup must find its row."""
    t = bench.profiled_traffic(CFG.name, prec, kname)
    assert t is not None and t > 1e9, (kname, prec, t)


def test_output_line_serialises_numpy_values():
    """The JSON line holds numpy scalars from the error and roofline sums
    (np.float64, np.bool_): they print as plain JSON values."""
    import json
    import numpy as np
    line = {"a": np.float64(1.5), "b": np.bool_(True), "c": {"d": np.int64(3)}, "e": np.zeros(2)}
    assert json.loads(json.dumps(line, default=bench.json_default)) == {"a": 1.5, "b": True, "c": {"d": 3},
                                                                       "e": [0.0, 0.0]}
