"""ctypes binding of libopentbl_amd.so (the C ABI in include/opentbl.h) and a
small host-side mirror of the reference's driver interface.

The library is loaded from this package directory; if it has not been built
(python -m opentbl_amd.build) every entry point raises, there is no fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libopentbl_amd.so")
_LIB = None

D, I, SZ = C.c_double, C.c_int, C.c_size_t
PD = C.POINTER(C.c_double)
VP = C.c_void_p

OTBL_FP64, OTBL_MIXED = 0, 1
OTBL_DT_EVERY_STEP, OTBL_DT_FROZEN = 0, 1
FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "rhsupa": 4, "rhsvpa": 5, "rhswpa": 6}
# values per diagnostics record (OTBL_DIAG_LEN) and their names in order
DIAG_LEN = 49
DIAG_NAMES = (["tiempo", "rthout", "utauout", "gamma1", "gamma2", "dout", "jtop", "ntop1", "ntop2"]
              + [f"{n}{i}" for n in ("x_plane", "u_inf", "u_tau", "d_out", "r_thout") for i in range(1, 9)])


class Params(C.Structure):
    _fields_ = [("nx", I), ("ny", I), ("nz", I), ("xout", I), ("re", D), ("ax", D),
                ("az", D), ("cfl", D), ("precision", I), ("dt_mode", I)]


# exported entry points: name -> (restype, argtypes)
SIGNATURES = {
    "otbl_create": (I, [C.POINTER(Params), PD, C.POINTER(VP)]),
    "otbl_destroy": (I, [VP]),
    "otbl_last_error": (C.c_char_p, []),
    "otbl_set_field": (I, [VP, I, PD]),
    "otbl_get_field": (I, [VP, I, PD]),
    "otbl_field_len": (SZ, [VP, I]),
    "otbl_set_inflow": (I, [VP, PD, PD, PD]),
    "otbl_get_um": (I, [VP, PD]),
    "otbl_get_genflu": (I, [VP, PD]),
    "otbl_get_ener": (I, [VP, PD]),
    "otbl_diag_enable": (I, [VP, I]),
    "otbl_diag_read": (I, [VP, PD, I, C.POINTER(I), C.POINTER(C.c_longlong)]),
    "otbl_retheta_cf": (I, [VP, PD, PD]),
    "otbl_set_time": (I, [VP, D, D, D]),
    "otbl_get_time": (I, [VP, PD, PD]),
    "otbl_get_coef": (I, [VP, C.c_char_p, PD, C.POINTER(I), C.POINTER(I)]),
    "otbl_rhsp": (I, [VP, I]),
    "otbl_outflow_correction": (I, [VP]),
    "otbl_pois": (I, [VP, D]),
    "otbl_pois_m": (I, [VP, I]),
    "otbl_substep": (I, [VP, I]),
    "otbl_step": (I, [VP, I]),
    "otbl_use_graphs": (I, [VP, I]),
    "otbl_sync": (I, [VP]),
    "otbl_interpxx": (I, [VP, VP, VP, I, I]),
    "otbl_difvisxx": (I, [VP, VP, VP, VP, I, I]),
    "otbl_interpyy": (I, [VP, VP, VP, I, I]),
    "otbl_differzy": (I, [VP, VP, VP, I, I]),
    "otbl_vistyy": (I, [VP, VP, VP, I, I]),
    "otbl_implzy": (I, [VP, VP, VP, I, D, I]),
    "otbl_fourxz": (I, [VP, VP, VP, I, I]),
    "otbl_cosftx": (I, [VP, VP, VP, I]),
    "otbl_cosft_lines": (I, [VP, VP, I, I]),
    "otbl_soltriy": (I, [VP, VP, VP, I, PD, I]),
    "otbl_soltriy_noalloc": (I, [VP, VP, VP, I, PD, I]),
    "otbl_coefinv": (I, [VP, PD, I, I]),
    "otbl_soltrix": (I, [VP, PD, I, I, I, PD, I, I, PD, C.POINTER(I)]),
    "otbl_coefb": (I, [I, I, C.c_double, C.c_double, C.c_double, C.c_double, PD]),
    "otbl_ensemble_create": (I, [C.POINTER(VP), I, C.POINTER(VP)]),
    "otbl_ensemble_step": (I, [VP, I]),
    "otbl_ensemble_sync": (I, [VP]),
    "otbl_ensemble_destroy": (I, [VP]),
    "otbl_dev_alloc": (I, [SZ, C.POINTER(VP)]),
    "otbl_dev_free": (I, [VP]),
    "otbl_h2d": (I, [VP, VP, SZ]),
    "otbl_d2h": (I, [VP, VP, SZ]),
    "otbl_d2d": (I, [VP, VP, SZ]),
    "otbl_get_paths": (I, [VP, C.c_char_p, SZ]),
    "otbl_autotune": (I, [VP, I, C.c_char_p, SZ]),
    "otbl_kprof": (I, [VP, I]),
    "otbl_kprof_report": (I, [VP, C.c_char_p, SZ]),
    "otbl_stats_name": (C.c_char_p, [I]),
    "otbl_stats_enable": (I, [VP, I]),
    "otbl_stats_schedule": (I, [VP, I, I]),
    "otbl_stats_request": (I, [VP]),
    "otbl_statsp": (I, [VP]),
    "otbl_stats_get": (I, [VP, I, PD]),
    "otbl_stats_ical": (I, [VP, C.POINTER(I)]),
    "otbl_stats_reset": (I, [VP]),
    "otbl_stats_get_ens": (I, [VP, C.POINTER(C.c_float)]),
}

# the accumulators of otbl_stats_* (OTBL_NSTAT), in the library's order
NSTAT = 40


class OtblError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OtblError(f"{LIB_PATH} is not built: run python -m opentbl_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = lib().otbl_last_error().decode()
        raise OtblError(f"{what}: status {rc}: {msg}")


def _pd(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


class DevArray:
    """A float64 device buffer managed through the C ABI."""

    def __init__(self, n):
        self.n = int(n)
        p = VP()
        check(lib().otbl_dev_alloc(self.n * 8, C.byref(p)), "dev_alloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a).view(np.float64).ravel()
        d = cls(a.size)
        check(lib().otbl_h2d(d.ptr, a.ctypes.data_as(VP), a.size * 8), "h2d")
        return d

    def to_host(self):
        out = np.empty(self.n, np.float64)
        check(lib().otbl_d2h(out.ctypes.data_as(VP), self.ptr, self.n * 8), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().otbl_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuBox:
    """One box of the MI355X substep, set up from an ic.make_state() dict."""

    def __init__(self, cfg, state, precision=OTBL_FP64, dt_mode=OTBL_DT_EVERY_STEP, graphs=False):
        L = lib()
        self.cfg = cfg
        prm = Params(cfg.nx, cfg.ny, cfg.nz, cfg.xout, state["re"], cfg.ax, cfg.az, cfg.cfl,
                     precision, dt_mode)
        y = np.ascontiguousarray(state["y"], np.float64)
        h = VP()
        check(L.otbl_create(C.byref(prm), _pd(y), C.byref(h)), "otbl_create")
        self.h = h
        for name in ("u", "v", "w", "p"):
            self.set(name, state[name])
        for name in ("rhsupa", "rhsvpa", "rhswpa"):
            if name in state:
                self.set(name, state[name])
        u0 = np.ascontiguousarray(state["u0"], np.float64)
        v0 = np.ascontiguousarray(state["v0"], np.float64)
        um = np.ascontiguousarray(state["um"], np.float64)
        check(L.otbl_set_inflow(self.h, _pd(u0), _pd(v0), _pd(um)), "set_inflow")
        check(L.otbl_set_time(self.h, state.get("tiempo", 0.0), state.get("timeinit", 0.0),
                              state.get("dt", 0.0)), "set_time")
        check(L.otbl_use_graphs(self.h, int(graphs)), "use_graphs")

    def close(self):
        if getattr(self, "h", None):
            lib().otbl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, name):
        c = self.cfg
        R = c.ny if name in ("v", "p", "rhsvpa") else c.ny + 1
        return (c.nx, R, c.mz)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, np.complex128).reshape(self._shape(name))
        check(lib().otbl_set_field(self.h, FIELDS[name], _pd(a.view(np.float64))), "set_field")

    def get(self, name):
        out = np.empty(self._shape(name), np.complex128)
        check(lib().otbl_get_field(self.h, FIELDS[name], _pd(out.view(np.float64))), "get_field")
        return out

    def rhsp(self, m):
        check(lib().otbl_rhsp(self.h, m), "rhsp")

    def outflow_correction(self):
        check(lib().otbl_outflow_correction(self.h), "outflow_correction")

    def pois(self, varstep):
        check(lib().otbl_pois(self.h, varstep), "pois")

    def pois_m(self, m):
        check(lib().otbl_pois_m(self.h, m), "pois_m")

    def substep(self, m):
        check(lib().otbl_substep(self.h, m), "substep")

    def step(self, n=1):
        check(lib().otbl_step(self.h, n), "step")

    def use_graphs(self, on):
        check(lib().otbl_use_graphs(self.h, int(on)), "use_graphs")

    def sync(self):
        check(lib().otbl_sync(self.h), "sync")

    def time(self):
        t, d = D(), D()
        check(lib().otbl_get_time(self.h, C.byref(t), C.byref(d)), "get_time")
        return t.value, d.value

    def um(self):
        o = np.zeros(self.cfg.ny + 1)
        check(lib().otbl_get_um(self.h, _pd(o)), "get_um")
        return o

    def genflu_info(self):
        o = np.zeros(6)
        check(lib().otbl_get_genflu(self.h, _pd(o)), "get_genflu")
        return dict(zip(("rthin", "din", "rthout", "dout", "gamma1", "gamma2"), o))

    def retheta_cf(self):
        """Re_theta(x) and Cf(x) at every plane (two arrays of nx)."""
        rth, cf = np.zeros(self.cfg.nx), np.zeros(self.cfg.nx)
        check(lib().otbl_retheta_cf(self.h, _pd(rth), _pd(cf)), "retheta_cf")
        return rth, cf

    def ener(self):
        o = np.zeros(15)
        check(lib().otbl_get_ener(self.h, _pd(o)), "get_ener")
        return o

    def diag_enable(self, capacity):
        """Per-substep info_per_step records into a device ring of `capacity`
        records (0: off)."""
        check(lib().otbl_diag_enable(self.h, int(capacity)), "diag_enable")

    def diag_read(self, max_records=4096):
        """The records written since the last read, oldest first, as an
        (n, DIAG_LEN) array, and the count lost to wrap-around."""
        out = np.zeros((max(1, max_records), DIAG_LEN))
        n, lost = I(), C.c_longlong()
        check(lib().otbl_diag_read(self.h, _pd(out), int(max_records), C.byref(n), C.byref(lost)), "diag_read")
        return out[:n.value].copy(), lost.value

    def coef(self, name):
        r, c = I(), I()
        check(lib().otbl_get_coef(self.h, name.encode(), None, C.byref(r), C.byref(c)), "get_coef")
        out = np.zeros(r.value * c.value)
        check(lib().otbl_get_coef(self.h, name.encode(), _pd(out), C.byref(r), C.byref(c)), "get_coef")
        return out.reshape(c.value, r.value).T if r.value > 1 else out

    def paths(self):
        """The kernel paths chosen at setup, e.g. {"y": "chunk20", "z": "engine1024", ...}."""
        buf = C.create_string_buffer(256)
        check(lib().otbl_get_paths(self.h, buf, len(buf)), "get_paths")
        return dict(kv.split(":") for kv in buf.value.decode().split())

    def autotune(self, reps=5):
        """Time the candidate launch shapes on this box and keep the fastest;
        returns the timings as text."""
        buf = C.create_string_buffer(4096)
        check(lib().otbl_autotune(self.h, int(reps), buf, len(buf)), "autotune")
        return buf.value.decode()

    def kprof(self, on):
        check(lib().otbl_kprof(self.h, int(on)), "kprof")

    def kprof_report(self):
        buf = C.create_string_buffer(1 << 16)
        check(lib().otbl_kprof_report(self.h, buf, len(buf)), "kprof_report")
        rows = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            rows[name] = (int(n), float(ms))
        return rows

    def op(self, name, *args):
        check(getattr(lib(), "otbl_" + name)(self.h, *args), name)

    # statistics (stotf.F90:51-291)
    def stats_enable(self, on=True):
        check(lib().otbl_stats_enable(self.h, int(on)), "stats_enable")

    def stats_schedule(self, avcrt, stest=0):
        """A sample at substep 1 of every step istep with istep % avcrt == 0
        and istep > stest (bl_1.F90:92)."""
        check(lib().otbl_stats_schedule(self.h, int(avcrt), int(stest)), "stats_schedule")

    def stats_request(self):
        check(lib().otbl_stats_request(self.h), "stats_request")

    def statsp(self):
        check(lib().otbl_statsp(self.h), "statsp")

    def stats(self):
        """The accumulators by name, each (nx, ny+1) (plane, row)."""
        out = {}
        for q in range(NSTAT):
            a = np.zeros(self.cfg.nx * (self.cfg.ny + 1))
            check(lib().otbl_stats_get(self.h, q, _pd(a)), "stats_get")
            out[lib().otbl_stats_name(q).decode()] = a.reshape(self.cfg.nx, self.cfg.ny + 1)
        return out

    def stats_ical(self):
        n = I()
        check(lib().otbl_stats_ical(self.h, C.byref(n)), "stats_ical")
        return n.value

    def stats_reset(self):
        check(lib().otbl_stats_reset(self.h), "stats_reset")

    def ens(self):
        """The enstrophy field of the last sample, (nx, ny, nz) float32."""
        c = self.cfg
        a = np.zeros(c.nx * c.ny * c.nz, np.float32)
        check(lib().otbl_stats_get_ens(self.h, a.ctypes.data_as(C.POINTER(C.c_float))), "stats_get_ens")
        return a.reshape(c.nx, c.ny, c.nz)


class Ensemble:
    """Independent boxes advanced together: each RK3 substep of all of them
    is one captured graph whose branches are the boxes' substeps.  The boxes
    (GpuBox objects) stay usable on their own; close the ensemble first."""

    def __init__(self, boxes):
        self.boxes = list(boxes)
        arr = (VP * len(self.boxes))(*[b.h for b in self.boxes])
        h = VP()
        check(lib().otbl_ensemble_create(arr, len(self.boxes), C.byref(h)), "ensemble_create")
        self.h = h

    def step(self, n=1):
        check(lib().otbl_ensemble_step(self.h, n), "ensemble_step")

    def sync(self):
        check(lib().otbl_ensemble_sync(self.h), "ensemble_sync")

    def close(self):
        if getattr(self, "h", None):
            lib().otbl_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
