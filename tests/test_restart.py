"""Restart files in the reference's HDF5 layout (opentbl_amd/restart.py over
libopentbl_io.so): escribezy/writeheader (escru.F90:159-251, 950-997) and
getstartzy/readheader (getstartzy-bg.F90:180-295, 490-581).

The layout is checked with HDF5's own h5dump, independently of the reader:
dataset names, types and shapes, and single elements of "value" at
(plane, row, real index) against the field they were written from.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from opentbl_amd import config, ic, restart

pytestmark = pytest.mark.skipif(not restart.available(), reason="libopentbl_io.so not built (no HDF5 C library)")

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def saved(tmp_path_factory):
    cfg = config.small(nx=20, ny=30, nz=24)
    st = ic.make_state(cfg)
    base = str(tmp_path_factory.mktemp("rst") / "bl.001")
    hdr = restart.header(cfg, st["re"], tiempo=12.5, dt=0.031, timeinit=2.0)
    um = np.asarray(st["um"]) * 1.001
    restart.save(base, hdr, st, st["y"], um)
    return cfg, st, base, um


def f32(a):
    a = np.asarray(a)
    return a.real.astype(np.float32).astype(np.float64) + 1j * a.imag.astype(np.float32).astype(np.float64)


def test_round_trip(saved):
    cfg, st, base, um = saved
    d = restart.load(base)
    h = d["header"]
    assert (h["nx"], h["ny"], h["nz2"], h["xout"], h["procs"]) == (cfg.nx, cfg.ny, cfg.nz2, cfg.xout, 1)
    assert h["tiempo"] == 12.5 and h["dt"] == 0.031 and h["timeinit"] == 2.0
    assert h["re"] == st["re"] and h["cfl"] == cfg.cfl
    assert h["lx"] == pytest.approx(cfg.Lx, rel=1e-15) and h["lz"] == pytest.approx(cfg.Lz, rel=1e-15)
    assert np.array_equal(d["y"], st["y"]) and np.array_equal(d["um"], um)
    assert d["variables"] == {"u": "u", "v": "v", "w": "w", "p": "p"}
    for n in ("u", "v", "w", "p"):
        # the fields travel as 4-byte reals (escru.F90:180-236)
        assert np.array_equal(d[n], f32(st[n])), n


def _h5dump(*args):
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_layout_seen_by_h5dump(saved):
    cfg, st, base, um = saved
    hdr = _h5dump("-H", restart.path(base, "v"))
    # value: 4-byte reals, C dims (nx, ny, nz1) = Fortran (nz1, ny, nx)
    m = re.search(r'DATASET "value" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)', hdr)
    assert m and m.group(1) == "H5T_IEEE_F32LE", hdr
    assert [int(x) for x in m.group(2).split(",")] == [cfg.nx, cfg.ny, cfg.nz1]
    for name in ("tiempo", "cfl", "Re", "lx", "ly", "lz", "timeinit", "dt"):
        m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)' % name, hdr)
        assert m and m.group(1) == "H5T_IEEE_F64LE" and m.group(2).strip() == "1", name
    for name in ("nx", "ny", "nz2", "xout", "procs"):
        m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)' % name, hdr)
        assert m and m.group(1) == "H5T_STD_I32LE" and m.group(2).strip() == "1", name
    for name, n in (("y", cfg.ny + 2), ("um", cfg.ny + 1)):
        m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)' % name, hdr)
        assert m and int(m.group(2)) == n, name
    assert 'DATASET "Variable"' in hdr
    # single elements: plane i, row j, real index kr (2k Re, 2k+1 Im of mode k)
    v = np.ascontiguousarray(st["v"]).view(np.float64)
    for i, j, kr in ((0, 0, 0), (3, 7, 5), (cfg.nx - 1, cfg.ny - 1, cfg.nz1 - 1), (11, 2, 2)):
        out = _h5dump("-d", "value", "-s", f"{i},{j},{kr}", "-c", "1,1,1", "-y", "-w", "0",
                      restart.path(base, "v"))
        body = out[out.index("DATA {"):]
        got = float(re.findall(r"[-+]?\d[\d.]*(?:e[-+]?\d+)?", body.split("{", 1)[1])[0])
        # h5dump prints 6 significant digits
        assert got == pytest.approx(float(np.float32(v[i, j, kr])), rel=1e-5, abs=1e-30), (i, j, kr)
    assert _h5dump("-d", "nz2", restart.path(base, "u")).count(f"{cfg.nz2}") >= 1


def test_other_spanwise_resolution(saved):
    """getstartzy copies the first min(nz1 file, nz1 run) reals of each row
    and leaves the rest zero (getstartzy-bg.F90:205-206, 230)."""
    cfg, st, base, um = saved
    wide = restart.load(base, nz1=cfg.nz1 + 8)
    narrow = restart.load(base, nz1=cfg.nz1 - 6)
    u = f32(st["u"]).view(np.float64).reshape(cfg.nx, cfg.ny + 1, cfg.nz1)
    uw = wide["u"].view(np.float64)
    un = narrow["u"].view(np.float64)
    assert np.array_equal(uw[..., :cfg.nz1], u) and not uw[..., cfg.nz1:].any()
    assert np.array_equal(un, u[..., :cfg.nz1 - 6])


def test_state_from_restart(saved):
    """A state read back: u0 and v0 the plane-1 mean profiles, um, tiempo,
    dt and Re from the header."""
    cfg, st, base, um = saved
    s2 = restart.load_state(base, cfg)
    assert np.array_equal(s2["u0"], f32(st["u"])[0, :, 0].real)
    assert np.array_equal(s2["v0"], f32(st["v"])[0, :, 0].real)
    assert np.array_equal(s2["um"], um) and s2["tiempo"] == 12.5 and s2["dt"] == 0.031
    with pytest.raises(restart.RestartError):
        restart.load_state(base, config.small(nx=22, ny=30, nz=24))


def test_errors(saved, tmp_path):
    cfg, st, base, um = saved
    with pytest.raises(restart.RestartError, match="cannot open"):
        restart.load(str(tmp_path / "missing"))
    f = np.zeros((cfg.nx, cfg.ny + 3, cfg.nz1))
    L = restart.lib()
    rc = L.otbl_io_read_field(restart.path(base, "u").encode(), f.ctypes.data_as(restart.PD), cfg.nx,
                              cfg.ny + 3, cfg.nz1)
    assert rc == 1 and b"rows per plane" in L.otbl_io_last_error()
