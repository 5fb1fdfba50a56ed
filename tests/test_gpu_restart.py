"""A GPU box written to restart files and read back into a new box
(escribezy / getstartzy, escru.F90:159-251, getstartzy-bg.F90:180-295): the
new box holds the first box's fields rounded to 4-byte reals (the restart
files' precision), its running mean um, tiempo and dt; stepping on from
both gives the same flow to the restart's FP32 rounding."""
import numpy as np
import pytest

from opentbl_amd import config, ic, restart
from opentbl_amd.lib import GpuBox

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not restart.available(), reason="libopentbl_io.so not built")]


def f32(a):
    return a.real.astype(np.float32).astype(np.float64) + 1j * a.imag.astype(np.float32).astype(np.float64)


def test_restart_round_trip_on_gpu(tmp_path):
    cfg = config.small(name="rst", nx=48, ny=40, nz=36, kzmax=6)
    st = ic.make_state(cfg)
    a = GpuBox(cfg, st, graphs=True)
    a.step(2)
    base = str(tmp_path / "bl.002")
    restart.save_box(base, a, st)
    s2 = restart.load_state(base, cfg)
    b = GpuBox(cfg, s2, graphs=True)
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(b.get(n), f32(a.get(n))), n
    assert np.array_equal(b.um(), a.um())
    assert b.time()[0] == a.time()[0]
    a.step(1)
    b.step(1)
    scale = max(np.abs(a.get(n)).max() for n in ("u", "v", "w"))
    err = max(np.abs(b.get(n) - a.get(n)).max() for n in ("u", "v", "w")) / scale
    a.close()
    b.close()
    print("one step after the restart: velocity difference", err, flush=True)
    assert err < 1e-6
