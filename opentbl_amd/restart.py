"""Restart files in the reference's HDF5 layout (libopentbl_io.so).

    save(base, cfg, fields, y, um, re, tiempo, dt, timeinit=0.0)
    load(base, nz1=None) -> dict(header, variable names, y, um, u, v, w, p, u0, v0)
    save_box(base, box, state) / state = load_state(base, cfg)

One file per field, <base>.u.h5, .v.h5, .w.h5, .p.h5, as escribezy writes
them (escru.F90:44-48, 159-251) and getstartzy reads them
(getstartzy-bg.F90:55-58, 180-295): the field as 4-byte reals in a Fortran
(nz1, R, nx) array, with the writeheader header (escru.F90:950-997).  A state
read back starts the run the way getstartzy leaves it: u0 and v0 are the
mean (mode 0) profiles of plane 1 (getstartzy-bg.F90:273-274), um, y, tiempo
and dt come from the header.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IOLIB = os.path.join(HERE, "libopentbl_io.so")
_LIB = None

PD = C.POINTER(C.c_double)


class Header(C.Structure):
    _fields_ = [("tiempo", C.c_double), ("cfl", C.c_double), ("re", C.c_double), ("lx", C.c_double),
                ("ly", C.c_double), ("lz", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz2", C.c_int),
                ("xout", C.c_int), ("procs", C.c_int), ("timeinit", C.c_double), ("dt", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class StatsHeader(C.Structure):
    _fields_ = [("tiempo", C.c_double), ("cfl", C.c_double), ("re", C.c_double), ("ax", C.c_double),
                ("ay", C.c_double), ("az", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz2", C.c_int),
                ("ical", C.c_int)]


class RestartError(RuntimeError):
    pass


def available():
    return os.path.exists(IOLIB)


def lib():
    global _LIB
    if _LIB is None:
        if not available():
            raise RestartError(f"{IOLIB} is not built (it needs the HDF5 C library): run "
                               "python -c 'from opentbl_amd import build; build.build_io()'")
        L = C.CDLL(IOLIB)
        L.otbl_io_write_field.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Header), PD, PD, PD, C.c_int]
        L.otbl_io_read_header.argtypes = [C.c_char_p, C.POINTER(Header), C.c_char_p, C.c_size_t, PD, PD]
        L.otbl_io_read_field.argtypes = [C.c_char_p, PD, C.c_int, C.c_int, C.c_int]
        L.otbl_io_write_ens.argtypes = [C.c_char_p, C.POINTER(Header), PD, PD, C.POINTER(C.c_float), C.c_int]
        L.otbl_io_write_stats.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(StatsHeader), PD, PD]
        L.otbl_io_last_error.restype = C.c_char_p
        for f in (L.otbl_io_write_field, L.otbl_io_read_header, L.otbl_io_read_field, L.otbl_io_write_ens,
                  L.otbl_io_write_stats):
            f.restype = C.c_int
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RestartError(f"{what}: status {rc}: {lib().otbl_io_last_error().decode()}")


def _pd(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


def _real(a):
    """A planes-layout field as float64 (nx, R, nz1): complex (nx, R, Mz)
    fields are viewed as their real pairs."""
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        a = np.ascontiguousarray(a, np.complex128).view(np.float64)
    return np.ascontiguousarray(a, np.float64)


def path(base, name):
    return f"{base}.{name}.h5"


def header(cfg, re, tiempo, dt, timeinit=0.0, procs=1):
    """The writeheader values of a box (escru.F90:172-174: lx = ax pi,
    ly = ay pi, lz = 2 pi az)."""
    return Header(tiempo=tiempo, cfl=cfg.cfl, re=re, lx=cfg.ax * math.pi, ly=cfg.Ly, lz=cfg.az * 2 * math.pi,
                  nx=cfg.nx, ny=cfg.ny, nz2=cfg.nz2, xout=cfg.xout, procs=procs, timeinit=timeinit, dt=dt)


def save(base, hdr, fields, y, um):
    """Write <base>.{u,v,w,p}.h5 from FP64 fields (complex (nx, R, Mz) or
    real (nx, R, nz1))."""
    y = np.ascontiguousarray(y, np.float64)
    um = np.ascontiguousarray(um, np.float64)
    if y.size != hdr.ny + 2 or um.size != hdr.ny + 1:
        raise RestartError("y must hold ny+2 values and um ny+1")
    for name in ("u", "v", "w", "p"):
        f = _real(fields[name])
        R = f.shape[1]
        if f.shape != (hdr.nx, R, 2 * (hdr.nz2 + 1)):
            raise RestartError(f"{name}: shape {f.shape} does not match the header")
        _check(lib().otbl_io_write_field(path(base, name).encode(), name.encode(), C.byref(hdr), _pd(y), _pd(um),
                                         _pd(f), R), f"write {name}")


def read_header(fname):
    h = Header()
    _check(lib().otbl_io_read_header(fname.encode(), C.byref(h), None, 0, None, None), "read header")
    var = C.create_string_buffer(64)
    y, um = np.zeros(h.ny + 2), np.zeros(h.ny + 1)
    _check(lib().otbl_io_read_header(fname.encode(), C.byref(h), var, len(var), _pd(y), _pd(um)), "read header")
    return h, var.value.decode(), y, um


def load(base, nz1=None):
    """Read <base>.{u,v,w,p}.h5.  The header comes from the u file as in
    getstartzy (getstartzy-bg.F90:183-184); fields are FP64 complex
    (nx, R, nz1/2) arrays with nz1 the caller's (default: the file's),
    modes past the file's zero."""
    h, var, y, um = read_header(path(base, "u"))
    nz1 = nz1 or 2 * (h.nz2 + 1)
    out = {"header": h.as_dict(), "variables": {}, "y": y, "um": um}
    for name in ("u", "v", "w", "p"):
        R = h.ny + 1 if name in ("u", "w") else h.ny
        f = np.zeros((h.nx, R, nz1))
        _check(lib().otbl_io_read_field(path(base, name).encode(), _pd(f), h.nx, R, nz1), f"read {name}")
        out[name] = f.view(np.complex128)
        out["variables"][name] = read_header(path(base, name))[1]
    return out


def save_box(base, box, state, timeinit=0.0):
    """Write the state of a GpuBox: its fields, running mean um, tiempo and
    dt; y and Re from the state it was set up with."""
    tiempo, dt = box.time()
    hdr = header(box.cfg, state["re"], tiempo, dt, timeinit)
    save(base, hdr, {n: box.get(n) for n in ("u", "v", "w", "p")}, state["y"], box.um())


def load_state(base, cfg):
    """A state dict for GpuBox / OracleBox from restart files written for
    the box cfg: fields, y, Re, um, tiempo, dt, timeinit, and u0, v0 from
    plane 1 (getstartzy-bg.F90:273-274)."""
    d = load(base, cfg.nz1)
    h = d["header"]
    if (h["nx"], h["ny"]) != (cfg.nx, cfg.ny):
        raise RestartError(f"restart is {h['nx']}x{h['ny']}, the box {cfg.nx}x{cfg.ny}")
    st = {n: d[n] for n in ("u", "v", "w", "p")}
    st.update(y=d["y"], re=h["re"], um=d["um"], tiempo=h["tiempo"], dt=h["dt"], timeinit=h["timeinit"],
              u0=st["u"][0, :, 0].real.copy(), v0=st["v"][0, :, 0].real.copy(), ax=cfg.ax, az=cfg.az,
              xout=cfg.xout)
    return st


# ---- statistics files (escrst, escru.F90:263-516) ----
# accumulator names in the library's order (otbl_stats_get ids)
STAT_NAMES = ("us", "vs", "ws", "uv", "ua", "va", "wa", "pp", "pm", "vortx", "vorty", "vortz", "vortxa",
              "vortya", "vortza", "uw", "vw", "dispu", "dispv", "dispw", "dispuv", "pup", "pvp", "pdudx",
              "pdvdy", "pdudy", "pdvdx", "pdwdz", "v3", "u2v", "v2u", "w2v", "dudx0", "dudy0", "dvdx0", "dvdy0",
              "dwdx0", "dwdy0", "u3", "w2u")
# per-plane record of the .st file: (name, rows beyond ny) (escru.F90:318-336)
ST_RECORD = (("us", 1), ("vs", 0), ("ws", 1), ("uv", 0), ("ua", 1), ("va", 0), ("wa", 1), ("pp", 0), ("pm", 0),
             ("vortx", 0), ("vorty", 0), ("vortz", 0), ("vortxa", 0), ("vortya", 0), ("vortza", 0), ("uw", 0),
             ("vw", 0))
# per-plane record of the .budget file (escru.F90:490-516)
BUDGET_RECORD = tuple((n, 0) for n in ("dispu", "dispv", "dispw", "dispuv", "pup", "pvp", "pdudx", "pdvdy", "pdudy",
                                       "pdvdx", "pdwdz", "v3", "u2v", "v2u", "w2v", "dudx0", "dudy0", "dvdx0",
                                       "dvdy0", "dwdx0", "dwdy0")) + (("u3", 1), ("w2u", 1))


def save_stats(st_path, budget_path, cfg, re, tiempo, ical, y, stats):
    """Write the .st and .budget statistics files (either path may be None)
    from a dict of accumulators, each (nx, ny+1) (plane, row) as
    GpuBox.stats() returns them.  The header carries ax, ay, az as bl_1
    passes them to escrst (domain lengths / pi, / pi, / 2 pi)."""
    arr = np.ascontiguousarray(np.stack([np.asarray(stats[n], np.float64) for n in STAT_NAMES]))
    if arr.shape != (len(STAT_NAMES), cfg.nx, cfg.ny + 1):
        raise RestartError(f"statistics of shape {arr.shape[1:]} for a {cfg.nx}x{cfg.ny} box")
    y = np.ascontiguousarray(y, np.float64)
    h = StatsHeader(tiempo=tiempo, cfl=cfg.cfl, re=re, ax=cfg.ax, ay=cfg.Ly / math.pi, az=cfg.az, nx=cfg.nx,
                    ny=cfg.ny, nz2=cfg.nz2, ical=int(ical))
    enc = lambda p: p.encode() if p else None  # noqa: E731
    _check(lib().otbl_io_write_stats(enc(st_path), enc(budget_path), C.byref(h), _pd(y), _pd(arr)), "write stats")


def _records(path):
    """The records of a big-endian Fortran sequential unformatted file."""
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        n = int.from_bytes(raw[o:o + 4], "big")
        out.append(raw[o + 4:o + 4 + n])
        if int.from_bytes(raw[o + 4 + n:o + 8 + n], "big") != n:
            raise RestartError(f"{path}: record markers differ at byte {o}")
        o += 8 + n
    return out


def read_stats(path, kind):
    """Read a .st (kind "st") or .budget (kind "budget") file: the header
    (tiempo, cfl, Re, ax, ay, az, nx, ny, nz2, ical), y(0:ny+1) and the
    accumulators it holds, each (nx, ny+1) with row ny+1 zero where the
    file stores ny rows."""
    rec = _records(path)
    hd = np.frombuffer(rec[0][:48], ">f8")
    nx, ny, nz2, ical = (int(v) for v in np.frombuffer(rec[0][48:64], ">i4"))
    head = dict(zip(("tiempo", "cfl", "re", "ax", "ay", "az"), hd.tolist()), nx=nx, ny=ny, nz2=nz2, ical=ical)
    y = np.frombuffer(rec[1], ">f8").astype(np.float64)
    layout, first = (ST_RECORD, 3) if kind == "st" else (BUDGET_RECORD, 2)
    out = {n: np.zeros((nx, ny + 1)) for n, _ in layout}
    for i in range(nx):
        vals = np.frombuffer(rec[first + i], ">f8")
        o = 0
        for n, extra in layout:
            out[n][i, :ny + extra] = vals[o:o + ny + extra]
            o += ny + extra
    return head, y, out


def save_ens(fname, hdr, y, um, ens):
    """The enstrophy file of escribezy (escru.F90:159-170): ens (nx, ny, nz)
    4-byte reals with the writeheader header, Variable "enstrophy"."""
    e = np.ascontiguousarray(ens, np.float32)
    y = np.ascontiguousarray(y, np.float64)
    um = np.ascontiguousarray(um, np.float64)
    _check(lib().otbl_io_write_ens(fname.encode(), C.byref(hdr), _pd(y), _pd(um),
                                   e.ctypes.data_as(C.POINTER(C.c_float)), e.shape[2]), "write ens")
