"""ctypes wrapper of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.
"""
import ctypes as C
import os
import subprocess
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.c_double
PD = C.POINTER(C.c_double)
I = C.c_int


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.oc_create.restype = C.c_void_p
        L.oc_create.argtypes = [I, I, I, I, D, D, D, D, PD, I, I]
        L.oc_destroy.argtypes = [C.c_void_p]
        L.oc_field_len.restype = C.c_size_t
        L.oc_field_len.argtypes = [C.c_void_p, I]
        L.oc_field.restype = PD
        L.oc_field.argtypes = [C.c_void_p, I]
        L.oc_set_inflow.argtypes = [C.c_void_p, PD, PD, PD]
        L.oc_get_um.argtypes = [C.c_void_p, PD]
        L.oc_set_time.argtypes = [C.c_void_p, D, D, D]
        L.oc_get_time.argtypes = [C.c_void_p, PD, PD]
        L.oc_get_genflu.argtypes = [C.c_void_p, PD]
        L.oc_get_ener.argtypes = [C.c_void_p, PD]
        L.oc_coef.restype = PD
        L.oc_coef.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(I), C.POINTER(I)]
        L.oc_scalar.restype = D
        L.oc_scalar.argtypes = [C.c_void_p, C.c_char_p]
        for n in ("oc_rhsp", "oc_substep"):
            getattr(L, n).argtypes = [C.c_void_p, I]
        L.oc_outflow_correction.argtypes = [C.c_void_p]
        L.oc_pois.argtypes = [C.c_void_p, D]
        L.oc_step.argtypes = [C.c_void_p]
        L.oc_op_interpxx.argtypes = [C.c_void_p, PD, PD, I, I]
        L.oc_op_difvisxx.argtypes = [C.c_void_p, PD, PD, PD, I, I]
        L.oc_op_interpyy.argtypes = [C.c_void_p, PD, PD, I, I]
        L.oc_op_differzy.argtypes = [C.c_void_p, PD, PD, I, I]
        L.oc_op_vistyy.argtypes = [C.c_void_p, PD, PD, I, I]
        L.oc_op_implzy.argtypes = [C.c_void_p, PD, PD, I, D, I]
        L.oc_c2r.argtypes = [C.c_void_p, PD, PD]
        L.oc_r2c.argtypes = [C.c_void_p, PD, PD]
        L.oc_cosft.argtypes = [I, I, PD, I]
        L.oc_retheta_cf.argtypes = [C.c_void_p, PD, PD]
        L.oc_max_threads.restype = I
        L.oc_max_threads.argtypes = []
        L.oc_coefb.argtypes = [I, I, D, D, D, D, PD]
        L.oc_statsp.argtypes = [C.c_void_p]
        L.oc_stat_name.restype = C.c_char_p
        L.oc_stat_name.argtypes = [I]
        L.oc_stats_get.argtypes = [C.c_void_p, I, PD]
        L.oc_stats_reset.argtypes = [C.c_void_p]
        L.oc_stats_ical.restype = I
        L.oc_stats_ical.argtypes = [C.c_void_p]
        L.oc_stats_ens.restype = C.POINTER(C.c_float)
        L.oc_stats_ens.argtypes = [C.c_void_p]
        L.oc_stats_coef.restype = PD
        L.oc_stats_coef.argtypes = [C.c_void_p, C.c_char_p]
        L.oc_stats_op.argtypes = [C.c_void_p, I, PD, PD, I]
        _LIB = L
    return _LIB


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


NSTAT = 40
FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "rhsupa": 4, "rhsvpa": 5, "rhswpa": 6}


class OracleBox:
    """One box of the oracle, initialised from an ic.make_state() dict."""

    def __init__(self, cfg, state, fp32_transposes=False, freeze_dt=False):
        L = lib()
        self.cfg = cfg
        y = np.ascontiguousarray(state["y"], np.float64)
        self.h = L.oc_create(cfg.nx, cfg.ny, cfg.nz, cfg.xout, state["re"], cfg.ax, cfg.az,
                             cfg.cfl, ptr(y), int(fp32_transposes), int(freeze_dt))
        for name in ("u", "v", "w", "p"):
            self.set(name, state[name])
        u0 = np.ascontiguousarray(state["u0"], np.float64)
        v0 = np.ascontiguousarray(state["v0"], np.float64)
        um = np.ascontiguousarray(state["um"], np.float64)
        L.oc_set_inflow(self.h, ptr(u0), ptr(v0), ptr(um))
        L.oc_set_time(self.h, state.get("tiempo", 0.0), state.get("timeinit", 0.0),
                      state.get("dt", 0.0))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    def _shape(self, name):
        c = self.cfg
        R = c.ny if name in ("v", "p", "rhsvpa") else c.ny + 1
        return (c.nx, R, c.mz)

    def view(self, name):
        """Complex (nx, R, Mz) view sharing the oracle's memory."""
        L = lib()
        n = L.oc_field_len(self.h, FIELDS[name])
        p = L.oc_field(self.h, FIELDS[name])
        a = np.ctypeslib.as_array(p, shape=(n,))
        return a.view(np.complex128).reshape(self._shape(name))

    def set(self, name, arr):
        self.view(name)[...] = arr

    def get(self, name):
        return self.view(name).copy()

    def coef(self, name):
        r, c = I(), I()
        p = lib().oc_coef(self.h, name.encode(), C.byref(r), C.byref(c))
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(r.value * c.value,)).copy()
        return a.reshape(c.value, r.value).T if r.value > 1 else a

    def scalar(self, name):
        return lib().oc_scalar(self.h, name.encode())

    def rhsp(self, m):
        lib().oc_rhsp(self.h, m)

    def outflow_correction(self):
        lib().oc_outflow_correction(self.h)

    def pois(self, varstep):
        lib().oc_pois(self.h, varstep)

    def substep(self, m):
        lib().oc_substep(self.h, m)

    def step(self):
        lib().oc_step(self.h)

    def time(self):
        t, d = D(), D()
        lib().oc_get_time(self.h, C.byref(t), C.byref(d))
        return t.value, d.value

    def um(self):
        o = np.zeros(self.cfg.ny + 1)
        lib().oc_get_um(self.h, ptr(o))
        return o

    def genflu_info(self):
        o = np.zeros(6)
        lib().oc_get_genflu(self.h, ptr(o))
        return dict(zip(("rthin", "din", "rthout", "dout", "gamma1", "gamma2"), o))

    def ener(self):
        o = np.zeros(15)
        lib().oc_get_ener(self.h, ptr(o))
        return o

    def retheta_cf(self):
        r = np.zeros(self.cfg.nx)
        c = np.zeros(self.cfg.nx)
        lib().oc_retheta_cf(self.h, ptr(r), ptr(c))
        return r, c

    # statistics (stotf.F90:51-291)
    def statsp(self):
        lib().oc_statsp(self.h)

    def stats(self):
        """The accumulators by name, each (nx, ny+1) (plane, row)."""
        L = lib()
        out = {}
        for q in range(NSTAT):
            a = np.zeros(self.cfg.nx * (self.cfg.ny + 1))
            L.oc_stats_get(self.h, q, ptr(a))
            out[L.oc_stat_name(q).decode()] = a.reshape(self.cfg.nx, self.cfg.ny + 1)
        return out

    def stats_reset(self):
        lib().oc_stats_reset(self.h)

    def stats_ical(self):
        return lib().oc_stats_ical(self.h)

    def ens(self):
        """The enstrophy field of the last sample, (nx, ny, nz) float32."""
        p = lib().oc_stats_ens(self.h)
        if not p:
            return None
        c = self.cfg
        return np.ctypeslib.as_array(p, shape=(c.nx * c.ny * c.nz,)).reshape(c.nx, c.ny, c.nz).copy()

    def stats_coef(self, name):
        """fd_dvdy, fd_dx, fd_ix, fd_iy, fd_iyp as (8, n) arrays."""
        p = lib().oc_stats_coef(self.h, name.encode())
        n = self.cfg.nx if name in ("fd_dx", "fd_ix") else self.cfg.ny
        return np.ctypeslib.as_array(p, shape=(8 * n,)).copy().reshape(n, 8).T

    def stats_op(self, op, a, n):
        """op 0 diffy, 1 interpy (u rows in), 2 interpy (p rows in) on n
        planes of (R, nz1) doubles; 3 diffx, 4 interpx on n pencils of nx."""
        a = np.ascontiguousarray(a, np.float64)
        c = self.cfg
        out = np.zeros(n * c.ny * c.nz1 if op <= 2 else n * c.nx)
        lib().oc_stats_op(self.h, op, ptr(a), ptr(out), n)
        return out

    # operator-level helpers on complex (nplanes, R, Mz) arrays
    def op(self, name, *args):
        return getattr(lib(), "oc_op_" + name)(self.h, *args)


def info_per_step_planes(u, y, re):
    """info_per_step (genflu.F90:247-295) at the eight sample planes, in
    numpy: rows x_plane, u_inf, u_tau, d_out, r_thout (5 x 8).

    u is a u-type field in the planes layout (nx, ny+1, nz1); the profile of
    a plane is its mode 0 (real part), um(j) = u[ip-1, j-1, 0].  The sample
    planes are ((nx-1) i)/8 + 1, i = 1..8: the first planes of ranks
    ((nummpi-1) i)/8 (genflu.F90:265) when every rank holds one plane.  y is
    y(0:ny+1); dy(j) = y(j+1) - y(j).  Sums run in the reference's order;
    0.99 is a default-real literal (genflu.F90:270)."""
    nx, R = u.shape[0], u.shape[1]
    ny = R - 1
    y = np.asarray(y, np.float64)
    dy = y[1:] - y[:-1]
    out = np.zeros((5, 8))
    for i in range(1, 9):
        ip = ((nx - 1) * i) // 8 + 1
        um = np.real(np.asarray(u[ip - 1, :, 0])).astype(np.float64)  # um(j) = um[j-1]
        uinf = 0.0
        for j in range(ny + 1 - 20, ny + 1 - 5 + 1):
            uinf += um[j - 1]
        uinf /= 16.0
        u99 = float(np.float32(0.99)) * uinf
        utau = np.sqrt(abs((um[1] - um[0]) / (y[2] - y[1]) / re))
        jtop = 0
        for j in range(1, ny - 1):
            if um[j - 1] >= u99:
                break
            jtop = j + 1
        s = 0.0
        for j in range(2, jtop + 4):
            s += (uinf - um[j - 1]) * um[j - 1] * dy[j - 1]
        out[:, i - 1] = (ip, uinf, utau, y[jtop], s * re / uinf)
    return out
