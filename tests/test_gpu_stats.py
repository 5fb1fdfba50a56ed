"""Statistics on the GPU (otbl_stats_*, stotf.F90:51-291 with -DNOCORR
-DNOSPECTRA): against the CPU FP64 oracle and the closed forms the oracle is
pinned by (tests/test_stats_oracle.py).

* the stats-only schemes the library derives (maximal-order moment
  conditions) equal the oracle's closed forms (coeft.f90:56-336);
* samples taken by otbl_statsp and by the sampling rule inside captured
  substep graphs (bl_1.F90:92) give the oracle's accumulators at 1e-10 and
  its enstrophy field at FP32 rounding, with the z sums on the LDS
  transforms (nz 24) and on the register-resident engine (nz 256);
* x-independent cubic profiles give the closed forms;
* a mixed box's statistics stay within 1e-5 of the FP64 box's;
* an ensemble whose boxes sample on the rule gives each box what the box
  gives alone, bit for bit;
* the statistics files written from the GPU accumulators read back equal.
"""
import numpy as np
import pytest

from opentbl_amd import config, ic, restart
from opentbl_amd.lib import Ensemble, GpuBox, OTBL_MIXED
from test_stats_oracle import closed_form_state, closed_form_stats

pytestmark = pytest.mark.gpu

STAT_TOL = 1e-10


def log(*a):
    print(*a, flush=True)


def stat_errors(g, o, ny):
    """max |g - o| / max |o| per accumulator over rows 1..ny (row ny+1 must
    be 0 in both)."""
    out = {}
    for n in o:
        assert np.all(g[n][:, ny] == 0.0), n
        s = max(np.abs(o[n][:, :ny]).max(), 1e-300)
        out[n] = float(np.abs(g[n][:, :ny] - o[n][:, :ny]).max() / s)
    return out


@pytest.fixture(scope="module", params=[24, 256], ids=["nz24-lds", "nz256-engine"])
def scfg(request):
    """nz 24 runs the z sums on the LDS transforms, nz 256 on the
    register-resident engine."""
    nz = request.param
    cfg = config.small(name=f"stg{nz}", nx=48, ny=40, nz=nz)
    return cfg, ic.make_state(cfg)


def test_stats_coefficients_vs_oracle(scfg, oracle_lib):
    cfg, st = scfg
    gb = GpuBox(cfg, st)
    ob = oracle_lib.OracleBox(cfg, st)
    for name in ("fd_dvdy", "fd_iy", "fd_iyp", "fd_dx", "fd_ix"):
        a, b = gb.coef(name), ob.stats_coef(name)
        assert a.shape == b.shape == (8, cfg.nx if name in ("fd_dx", "fd_ix") else cfg.ny), name
        used = np.ones(a.shape, bool)
        if name in ("fd_iy", "fd_iyp"):
            used[6, 0] = False          # the wall row is explicit: its pivot slot is never read
        if name == "fd_iyp":
            used[3, cfg.ny - 2:] = False  # rows ny-1, ny of p use 3 weights
        err = np.abs(a - b)[used].max() / np.abs(b[used]).max()
        log(name, "library vs oracle", err)
        assert err < 1e-12, name
    gb.close()


def test_stats_vs_oracle(scfg, oracle_lib):
    """One sample by otbl_statsp, then two steps with a sample at substep 1
    of each (avcrt = 1) inside captured graphs, against the oracle taking its
    samples before each step."""
    cfg, st = scfg
    gb = GpuBox(cfg, st, graphs=True)
    ob = oracle_lib.OracleBox(cfg, st)
    gb.stats_enable()
    gb.statsp()
    ob.statsp()
    gb.stats_schedule(1)
    for _ in range(2):
        ob.statsp()
        ob.step()
    gb.step(2)
    errs = stat_errors(gb.stats(), ob.stats(), cfg.ny)
    worst = max(errs, key=errs.get)
    log("stats vs oracle after 3 samples: worst", worst, errs[worst])
    assert gb.stats_ical() == ob.stats_ical() == 3
    assert all(v < STAT_TOL for v in errs.values()), errs
    e, eo = gb.ens(), ob.ens()
    assert np.abs(e - eo).max() <= 2e-7 * np.abs(eo).max()
    # the state is what the same steps give without statistics
    gr = GpuBox(cfg, st, graphs=True)
    gr.step(2)
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(gb.get(n), gr.get(n)), n
    gb.stats_reset()
    assert gb.stats_ical() == 0 and all(np.all(a == 0.0) for a in gb.stats().values())
    gb.close()
    gr.close()


@pytest.mark.parametrize("nz", [24, 256])
def test_stats_closed_form_gpu(nz):
    cfg = config.small(name=f"stc{nz}", nx=24, ny=40, nz=nz)
    st, prof, dprof = closed_form_state(cfg, ic.make_state(cfg))
    gb = GpuBox(cfg, st)
    gb.stats_enable()
    gb.statsp()
    got = gb.stats()
    gb.close()
    want = closed_form_stats(cfg, st, prof, dprof, float(np.float32(1.0 / cfg.az)))
    for n, wv in want.items():
        assert np.abs(got[n][:, :cfg.ny] - wv[None, :]).max() <= 1e-11 * max(1.0, np.abs(wv).max()), n


@pytest.mark.parametrize("nz", [256, 96])
def test_stats_mixed_vs_fp64(nz):
    """The mixed box samples its FP32 fields, with the mean-flow u lines in
    FP64 as the substep keeps them (register-engine and LDS z lengths):
    every accumulator within 1e-5 of the FP64 box's (relative to its largest
    value), the enstrophy within 1e-5.  The x derivatives of the mean v and w
    (dvdx0, dwdx0) are a small difference of FP32 values: their error is
    ~2e-4 of their own small size and is held to 1e-7 of the mean shear
    (max |dudy0|) instead."""
    cfg = config.small(name=f"stm{nz}", nx=130, ny=40, nz=nz, Lx=8.0 * 130 / 32, kzmax=8)
    st = ic.make_state(cfg)
    gd = GpuBox(cfg, st, graphs=True)
    gm = GpuBox(cfg, st, precision=OTBL_MIXED, graphs=True)
    for b in (gd, gm):
        b.stats_enable()
        b.stats_schedule(1)
        b.step(1)
    sm, sd = gm.stats(), gd.stats()
    errs = stat_errors(sm, sd, cfg.ny)
    log(f"mixed vs FP64 statistics (nz {nz}):", sorted(errs.items(), key=lambda kv: -kv[1])[:6])
    e = np.abs(gm.ens() - gd.ens()).max() / np.abs(gd.ens()).max()
    gd.close()
    gm.close()
    shear = np.abs(sd["dudy0"]).max()
    for n in ("dvdx0", "dwdx0"):
        d = np.abs(sm[n] - sd[n]).max() / shear
        log(f"  {n}: {errs.pop(n):.2e} of its own size, {d:.2e} of the mean shear")
        assert d < 1e-7, n
    assert all(v < 1e-5 for v in errs.values()), errs
    assert e < 1e-5


def test_stats_in_ensemble_match_boxes_alone():
    cfgs = [config.small(name="se0"), config.small(name="se1", nx=48, ny=56, nz=36, kzmax=6, seed=3)]
    states = [ic.make_state(c) for c in cfgs]
    eb = [GpuBox(c, s, graphs=True) for c, s in zip(cfgs, states)]
    al = [GpuBox(c, s, graphs=True) for c, s in zip(cfgs, states)]
    for b in eb + al:
        b.stats_enable()
    # box 0 samples every step, box 1 every second step after step 1
    for bs in (eb, al):
        bs[0].stats_schedule(1)
        bs[1].stats_schedule(2, 1)
    ens = Ensemble(eb)
    ens.step(4)
    ens.sync()
    for b in al:
        b.step(4)
    for a, b in zip(eb, al):
        assert a.stats_ical() == b.stats_ical()
        sa, sb = a.stats(), b.stats()
        for n in sa:
            assert np.array_equal(sa[n], sb[n]), n
        assert np.array_equal(a.ens(), b.ens())
        for n in ("u", "v", "w", "p"):
            assert np.array_equal(a.get(n), b.get(n)), n
    assert [b.stats_ical() for b in al] == [4, 2]   # box 1: steps 2 and 4
    ens.close()
    for b in eb + al:
        b.close()


@pytest.mark.skipif(not restart.available(), reason="libopentbl_io.so not built")
def test_stats_files_from_gpu(scfg, tmp_path):
    cfg, st = scfg
    gb = GpuBox(cfg, st)
    gb.stats_enable()
    gb.statsp()
    gb.statsp()
    s = gb.stats()
    ical = gb.stats_ical()
    t, _ = gb.time()
    ens = gb.ens()
    um = gb.um()
    gb.close()
    p1, p2 = str(tmp_path / "bl.01.000.st"), str(tmp_path / "bl.01.000.budget")
    restart.save_stats(p1, p2, cfg, st["re"], t, ical, st["y"], s)
    h, y, a = restart.read_stats(p1, "st")
    h2, _, b = restart.read_stats(p2, "budget")
    assert h["ical"] == h2["ical"] == 2 and h["nx"] == cfg.nx and h["ny"] == cfg.ny
    assert np.array_equal(y, st["y"])
    # row ny+1 of the accumulators is never written, so the rows the files
    # leave out read back as the zeros they were
    for d in (a, b):
        for n, v in d.items():
            assert np.array_equal(v, s[n]), n
    hdr = restart.header(cfg, st["re"], t, 0.0)
    restart.save_ens(str(tmp_path / "bl.001.ens.h5"), hdr, st["y"], um, ens)
