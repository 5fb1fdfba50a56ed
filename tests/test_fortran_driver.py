"""The Fortran driver (opentbl_amd/fortran/main.F90) reaches the GPU kernels
through the ISO_C_BINDING layer with the reference's rhsp / outflow_correction
/ pois interfaces; two steps of it match the CPU FP64 oracle, with the state
resident on the device and with the drop-ins mirroring the caller's arrays."""
import os
import subprocess

import numpy as np
import pytest

from opentbl_amd import build, config, ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_input(path, cfg, st, dt_mode=0):
    with open(path, "wb") as f:
        np.array([cfg.nx, cfg.ny, cfg.nz, cfg.xout, dt_mode], np.int32).tofile(f)
        np.array([st["re"], cfg.ax, cfg.az, cfg.cfl], np.float64).tofile(f)
        for a in (st["y"], st["u0"], st["v0"], st["um"]):
            np.ascontiguousarray(a, np.float64).tofile(f)
        for n in ("u", "v", "w", "p"):
            np.ascontiguousarray(st[n]).view(np.float64).tofile(f)


def read_output(path, cfg):
    raw = np.fromfile(path, np.float64)
    t, dt = raw[0], raw[1]
    off = 2
    out = {}
    for n, R in (("u", cfg.ny + 1), ("v", cfg.ny), ("w", cfg.ny + 1), ("p", cfg.ny)):
        cnt = cfg.nx * R * cfg.nz1
        out[n] = raw[off:off + cnt].view(np.complex128).reshape(cfg.nx, R, cfg.mz)
        off += cnt
    return t, dt, out


def test_fortran_driver_builds():
    if build.flang() is None:
        pytest.skip("no Fortran compiler")
    path = build.build_fortran()
    assert os.path.exists(path)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["resident", "mirror"])
def test_fortran_driver_matches_oracle(tmp_path, oracle_lib, mode):
    """resident: the state stays on the device between the drop-in calls;
    mirror: the drop-ins copy the fields in and out of the driver's arrays
    at every call, as the reference's callees update them in place."""
    drv = build.DRIVER
    if not os.path.exists(drv):
        pytest.skip("Fortran driver not built (no Fortran compiler where the build ran)")
    cfg = config.small()
    st = ic.make_state(cfg)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, cfg, st)
    r = subprocess.run([drv, fin, fout, "2"] + (["mirror"] if mode == "mirror" else []), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    t, dt, out = read_output(fout, cfg)
    ob = oracle_lib.OracleBox(cfg, st)
    ob.step()
    ob.step()
    to, dto = ob.time()
    assert abs(t - to) < 1e-12 * to and abs(dt - dto) < 1e-12 * dto
    for n in ("u", "v", "w", "p"):
        ref = ob.get(n)
        err = np.abs(out[n] - ref).max() / np.abs(ref).max()
        assert err < 1e-10, (n, err)
