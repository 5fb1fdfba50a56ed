"""Per-substep diagnostics records (info_per_step, genflu.F90:247-295) from
the device ring: the sample-plane columns against the oracle's numpy
restatement applied to the oracle's state before each substep, the
recycling columns against the oracle's genflu after that substep, graphs
against eager launches, the ring's wrap-around, the mixed substep against
the FP64 one, and switching the diagnostics on after graphs were
captured."""
import numpy as np
import pytest

from opentbl_amd import config, ic
from opentbl_amd.lib import DIAG_LEN, GpuBox, OTBL_MIXED, OtblError
from oracle.oracle import info_per_step_planes

pytestmark = pytest.mark.gpu

TOL = 1e-11


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def dbox():
    cfg = config.small(name="diag", nx=96, ny=40, nz=24, Lx=8.0 * 96 / 32)
    return cfg, ic.make_state(cfg)


def planes(rec):
    """x_plane, u_inf, u_tau, d_out, r_thout rows (5 x 8) of a record"""
    return rec[9:].reshape(5, 8)


def test_diag_vs_oracle(dbox, oracle_lib):
    cfg, st = dbox
    ob = oracle_lib.OracleBox(cfg, st)
    gb = GpuBox(cfg, st)
    gb.diag_enable(8)
    gb.step(1)
    recs, lost = gb.diag_read()
    y = np.asarray(st["y"])
    t0 = gb.time()
    gb.close()
    assert recs.shape == (3, DIAG_LEN) and lost == 0
    for m in (1, 2, 3):
        rec = recs[m - 1]
        want = info_per_step_planes(ob.get("u"), y, st["re"])  # the state this substep starts from
        ob.substep(m)
        got = planes(rec)
        assert np.array_equal(got[0], want[0]), (m, got[0], want[0])
        for r in range(1, 5):
            assert rel(got[r], want[r]) < TOL, (m, r, got[r], want[r])
        info = ob.genflu_info()
        assert rel(rec[1], info["rthout"]) < TOL and rel(rec[5], info["dout"]) < TOL
        assert rel(rec[3:5], [info["gamma1"], info["gamma2"]]) < TOL
        um = ob.um()
        assert rel(rec[2], np.sqrt(abs((um[1] - um[0]) / (y[2] - y[1]) / st["re"]))) < 1e-10
        assert rec[5] == y[int(rec[6])] and rec[7] >= 4 and rec[8] >= 4
    # tiempo advances at the end of a step (bl_1.F90:127): the same for its substeps
    assert recs[0, 0] == recs[1, 0] == recs[2, 0] and recs[0, 0] < t0[0]


def test_diag_graphs_match_eager(dbox):
    cfg, st = dbox
    ge = GpuBox(cfg, st)
    gg = GpuBox(cfg, st, graphs=True)
    for g in (ge, gg):
        g.diag_enable(16)
        g.step(2)
    re_, le = ge.diag_read()
    rg, lg = gg.diag_read()
    ge.close()
    gg.close()
    assert re_.shape == (6, DIAG_LEN) and le == lg == 0
    assert np.array_equal(re_, rg)


def test_diag_ring_wrap(dbox):
    """A ring of 2 records after 3 substeps keeps the last two and reports
    one lost; a partial read leaves the rest for the next one."""
    cfg, st = dbox
    ga = GpuBox(cfg, st)
    gb = GpuBox(cfg, st, graphs=True)
    ga.diag_enable(16)
    gb.diag_enable(2)
    ga.step(1)
    gb.step(1)
    full, _ = ga.diag_read()
    recs, lost = gb.diag_read()
    assert lost == 1 and np.array_equal(recs, full[1:])
    ga.step(1)
    gb.step(1)
    full, _ = ga.diag_read()
    first, lost = gb.diag_read(max_records=1)
    rest, lost2 = gb.diag_read()
    ga.close()
    gb.close()
    assert lost == 1 and lost2 == 0
    assert np.array_equal(first, full[1:2]) and np.array_equal(rest, full[2:3])


def test_diag_mixed_vs_fp64():
    cfg = config.small(name="diagmix", nx=130, ny=40, nz=256, Lx=8.0 * 130 / 32, kzmax=40)
    st = ic.make_state(cfg)
    gd = GpuBox(cfg, st)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    for g in (gd, gm):
        g.diag_enable(8)
        g.step(1)
    rd, _ = gd.diag_read()
    rm, _ = gm.diag_read()
    gd.close()
    gm.close()
    assert rd.shape == rm.shape == (3, DIAG_LEN)
    assert np.array_equal(rd[:, 9:17], rm[:, 9:17])  # sample planes
    assert rel(rm[:, 17:], rd[:, 17:]) < 1e-6
    assert rel(rm[:, 1:6], rd[:, 1:6]) < 1e-6


def test_diag_off_and_late_enable(dbox):
    cfg, st = dbox
    gb = GpuBox(cfg, st, graphs=True)
    with pytest.raises(OtblError, match="diagnostics are off"):
        gb.diag_read()
    gb.step(1)  # graphs captured without the diagnostics kernel
    gb.diag_enable(4)
    gb.step(1)
    recs, lost = gb.diag_read()
    gb.diag_enable(0)
    gb.step(1)
    gb.close()
    assert recs.shape == (3, DIAG_LEN) and lost == 0 and np.all(np.isfinite(recs))
