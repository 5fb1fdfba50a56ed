"""Statistics files in the reference's layouts (libopentbl_io.so through
opentbl_amd/restart.py): the .st and .budget files of escrst (Fortran
sequential unformatted records, big-endian: escru.F90:263-341, 461-516)
and the enstrophy HDF5 file of escribezy (escru.F90:159-170).

The records are parsed here byte by byte, independently of
restart.read_stats: record markers, header values and types, the order of
the arrays inside each plane's record.  The enstrophy file is checked with
HDF5's own h5dump.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from opentbl_amd import config, ic, restart

pytestmark = pytest.mark.skipif(not restart.available(), reason="libopentbl_io.so not built (no HDF5 C library)")

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    cfg = config.small(nx=20, ny=30, nz=24)
    st = ic.make_state(cfg)
    rng = np.random.default_rng(3)
    stats = {n: rng.standard_normal((cfg.nx, cfg.ny + 1)) for n in restart.STAT_NAMES}
    d = tmp_path_factory.mktemp("stats")
    p1, p2 = str(d / "bl.01.000.st"), str(d / "bl.01.000.budget")
    restart.save_stats(p1, p2, cfg, st["re"], 7.25, 11, st["y"], stats)
    return cfg, st, stats, p1, p2


def records(path):
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        (n,) = struct.unpack(">i", raw[o:o + 4])
        (m,) = struct.unpack(">i", raw[o + 4 + n:o + 8 + n])
        assert m == n
        out.append(raw[o + 4:o + 4 + n])
        o += 8 + n
    assert o == len(raw)
    return out


def check_header(rec, cfg, st):
    assert len(rec[0]) == 6 * 8 + 4 * 4
    t, cfl, re, ax, ay, az = struct.unpack(">6d", rec[0][:48])
    assert (t, cfl, re) == (7.25, cfg.cfl, st["re"])
    assert ax == cfg.ax and az == cfg.az and ay == pytest.approx(cfg.Ly / np.pi, rel=1e-15)
    assert struct.unpack(">4i", rec[0][48:]) == (cfg.nx, cfg.ny, cfg.nz2, 11)
    assert np.array_equal(np.frombuffer(rec[1], ">f8"), st["y"])


def test_st_file_layout(written):
    cfg, st, s, p1, _ = written
    rec = records(p1)
    ny = cfg.ny
    assert len(rec) == 3 + cfg.nx
    check_header(rec, cfg, st)
    assert struct.unpack(">d", rec[2]) == (0.0,)
    order = [("us", ny + 1), ("vs", ny), ("ws", ny + 1), ("uv", ny), ("ua", ny + 1), ("va", ny), ("wa", ny + 1),
             ("pp", ny), ("pm", ny), ("vortx", ny), ("vorty", ny), ("vortz", ny), ("vortxa", ny), ("vortya", ny),
             ("vortza", ny), ("uw", ny), ("vw", ny)]
    for i in range(cfg.nx):
        v = np.frombuffer(rec[3 + i], ">f8")
        assert v.size == sum(n for _, n in order)
        o = 0
        for name, n in order:
            assert np.array_equal(v[o:o + n], s[name][i, :n]), (i, name)
            o += n


def test_budget_file_layout(written):
    cfg, st, s, _, p2 = written
    rec = records(p2)
    ny = cfg.ny
    assert len(rec) == 2 + cfg.nx
    check_header(rec, cfg, st)
    cols = ["dispu", "dispv", "dispw", "dispuv", "pup", "pvp", "pdudx", "pdvdy", "pdudy", "pdvdx", "pdwdz", "v3",
            "u2v", "v2u", "w2v", "dudx0", "dudy0", "dvdx0", "dvdy0", "dwdx0", "dwdy0"]
    for i in range(cfg.nx):
        v = np.frombuffer(rec[2 + i], ">f8")
        assert v.size == 21 * ny + 2 * (ny + 1)
        bud = v[:21 * ny].reshape(21, ny)     # buf_bud(ny, 21), column-major
        for c, name in enumerate(cols):
            assert np.array_equal(bud[c], s[name][i, :ny]), (i, name)
        assert np.array_equal(v[21 * ny:21 * ny + ny + 1], s["u3"][i])
        assert np.array_equal(v[21 * ny + ny + 1:], s["w2u"][i])


def test_read_stats_round_trip(written):
    cfg, st, s, p1, p2 = written
    h, y, a = restart.read_stats(p1, "st")
    _, _, b = restart.read_stats(p2, "budget")
    assert h["ical"] == 11 and np.array_equal(y, st["y"])
    for d in (a, b):
        for n, v in d.items():
            R = v.shape[1] if np.any(v[:, cfg.ny]) else cfg.ny
            assert np.array_equal(v[:, :R], s[n][:, :R]), n


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_enstrophy_file_seen_by_h5dump(tmp_path):
    cfg = config.small(nx=16, ny=26, nz=12, kzmax=2)
    st = ic.make_state(cfg)
    ens = np.random.default_rng(5).random((cfg.nx, cfg.ny, cfg.nz)).astype(np.float32)
    f = str(tmp_path / "bl.001.ens.h5")
    hdr = restart.header(cfg, st["re"], 1.5, 0.01)
    restart.save_ens(f, hdr, st["y"], st["um"], ens)
    out = subprocess.run([H5DUMP, "-H", "-d", "value", f], capture_output=True, text=True, check=True).stdout
    assert f"( {cfg.nx}, {cfg.ny}, {cfg.nz} )" in out and "H5T_IEEE_F32" in out
    out = subprocess.run([H5DUMP, "-d", "Variable", f], capture_output=True, text=True, check=True).stdout
    assert '"enstrophy"' in out
    # elements (plane, row, point) by their C-order index; h5dump prints 6
    # significant digits
    for i, j, k in ((3, 5, 7), (15, 25, 11), (0, 0, 0)):
        out = subprocess.run([H5DUMP, "-d", f"value[{i},{j},{k};;1,1,1]", f], capture_output=True, text=True,
                             check=True).stdout
        val = float(out.split(f"({i},{j},{k}):")[1].split()[0])
        assert val == pytest.approx(float(ens[i, j, k]), rel=1e-5)
