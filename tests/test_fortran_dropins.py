"""The Fortran drop-ins of opentbl_amd/fortran/otbl_gpu.f90, called with the
reference's argument lists and array layouts (lip.f90, ffts.f90, change.f90,
pois.f90:211, coeft.f90:1244,1338,1360,1926), against the CPU FP64 oracle
or dense solves.

The binding is loaded as libotbl_fortran.so: the box is opened through its
C-callable otbl_f_open, the drop-ins are called by their Fortran names
(arguments by reference).  Pencil arrays are (nx, nlin) with line
p = kr + nz1 (j - 1); planes are (nz1, R, nx); complex (0:nz2) rows are nz1
reals.
"""
import ctypes as C
import os

import numpy as np
import pytest

from opentbl_amd import build, config, ic

DROPINS = ["rhsp_", "outflow_correction_", "pois_", "chp2x_", "chx2p_", "interpxx_", "difvisxx_", "interpyy_",
           "differzy_", "vistyy_", "implzy_", "fourxz_", "rfti_", "cosfti_", "cosftx_", "soltriy_", "coefinv_",
           "coef_", "mtr_", "statsp_", "soltrix_", "coefb_", "pointersfft_", "pointers_p2p_", "comm_setup_",
           "otbl_f_open", "otbl_f_close", "otbl_f_mirror", "otbl_f_box"]

OP_TOL = 1e-12


def flib():
    if not os.path.exists(build.FLIB):
        pytest.skip("Fortran binding not built (no Fortran compiler where the build ran)")
    return C.CDLL(build.FLIB)


def test_fortran_binding_exports_every_dropin():
    L = flib()
    missing = [n for n in DROPINS if not hasattr(L, n)]
    assert not missing, missing


def iref(v):
    return C.byref(C.c_int(int(v)))


def dref(v):
    return C.byref(C.c_double(float(v)))


def p(a):
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def planes(cfg, R, seed):
    """A random planes-layout field as a C-ordered (nx, R, nz1) array (the
    Fortran (nz1, R, nx) memory)."""
    return np.ascontiguousarray(np.random.default_rng(seed).standard_normal((cfg.nx, R, cfg.nz1)))


def to_pencils(a):
    """(nx, R, nz1) planes -> the Fortran pencil array pen(nx, nlin) as a
    C-ordered (nlin, nx) array."""
    nx = a.shape[0]
    return np.ascontiguousarray(a.reshape(nx, -1).T)


def from_pencils(pen, shape):
    return np.ascontiguousarray(pen.T.reshape(shape))


@pytest.fixture(scope="module")
def fbox(oracle_lib):
    L = flib()
    cfg = config.small()
    st = ic.make_state(cfg)
    y = np.ascontiguousarray(st["y"], np.float64)
    L.otbl_f_open.argtypes = [C.c_int] * 4 + [C.c_double] * 4 + [C.POINTER(C.c_double), C.c_int]
    L.otbl_f_open(cfg.nx, cfg.ny, cfg.nz, cfg.xout, st["re"], cfg.ax, cfg.az, cfg.cfl, p(y), 0)
    ob = oracle_lib.OracleBox(cfg, st)
    yield L, cfg, st, ob
    L.otbl_f_close()


pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("loguv", [1, 0])
def test_interpxx_dropin(fbox, loguv):
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    R = cfg.ny + 1 if loguv else cfg.ny
    g = planes(cfg, R, 1)
    ref = np.zeros_like(g)
    ob.op("interpxx", ptr(g), ptr(ref), R, loguv)
    pen = to_pencils(g)
    out = np.zeros_like(pen)
    dummy = np.zeros(16)
    L.interpxx_(p(pen), p(out), p(dummy), p(dummy), p(dummy), iref(2 if loguv else 1), iref(cfg.nz1 * R), iref(loguv))
    got = from_pencils(out, g.shape)
    assert np.abs(got - ref).max() < OP_TOL * np.abs(ref).max()


@pytest.mark.parametrize("vel", [1, 0])
def test_difvisxx_dropin_in_place(fbox, vel):
    """Called as rhsp.F90:182-192 does: the result overwrites the product."""
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    R = cfg.ny + 1
    g, q = planes(cfg, R, 2), planes(cfg, R, 3)
    ref = np.zeros_like(g)
    ob.op("difvisxx", ptr(g), ptr(q), ptr(ref), R, vel)
    pg, pq = to_pencils(g), to_pencils(q)
    dummy = np.zeros(32)
    L.difvisxx_(p(pg), p(pq), p(pg), p(dummy), p(dummy), p(dummy), p(dummy), p(dummy), p(dummy),
                iref(1 if vel else 2), iref(cfg.nz1 * R), iref(vel), dref(-1.0 / st["re"]))
    got = from_pencils(pg, g.shape)
    assert np.abs(got - ref).max() < OP_TOL * np.abs(ref).max()


@pytest.mark.parametrize("loguv", [0, 1])
@pytest.mark.parametrize("physical", [False, True], ids=["spectral", "physical"])
def test_interpyy_dropin(fbox, loguv, physical):
    """One plane of rows (dim, jee): spectral rows (dim = m = nz1) as in
    rhsp.F90:367,393 and physical rows (dim = nz + 2, m = nz) as in
    rhsp.F90:146, in place."""
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    jee = cfg.ny if loguv else cfg.ny + 1
    dim, m = (cfg.nz + 2, cfg.nz) if physical else (cfg.nz1, cfg.nz1)
    rng = np.random.default_rng(4 + loguv)
    rows = np.zeros((cfg.ny + 1, dim))
    rows[:jee, :m] = rng.standard_normal((jee, m))
    # the oracle on columns of nz1: pad the m columns into planes
    nb = (m + cfg.nz1 - 1) // cfg.nz1
    Rin = jee
    g = np.zeros((cfg.nx, Rin, cfg.nz1))
    for q in range(nb):
        c = rows[:jee, q * cfg.nz1:min(m, (q + 1) * cfg.nz1)]
        g[q, :, :c.shape[1]] = c
    ref = np.zeros((cfg.nx, cfg.ny + 1, cfg.nz1))
    ob.op("interpyy", ptr(g), ptr(ref), loguv, nb)
    want = np.concatenate([ref[q, :cfg.ny] for q in range(nb)], axis=1)[:, :m]
    buf = np.ascontiguousarray(rows)
    dummy = np.zeros(16)
    L.interpyy_(p(buf), p(buf), p(dummy), p(dummy), p(dummy), iref(jee), iref(loguv), iref(dim), iref(m))
    assert np.abs(buf[:cfg.ny, :m] - want).max() < OP_TOL * np.abs(want).max()


@pytest.mark.parametrize("form", [(3, -1, 0, "u"), (1, 0, 1, "v")], ids=["v_to_u", "u_to_v"])
def test_differzy_dropin(fbox, form):
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    chv, f1, f2, kind = form
    jee = cfg.ny + 1 if kind == "u" else cfg.ny
    which = 0 if kind == "u" else 1
    g = np.zeros((cfg.nx, cfg.ny + 1, cfg.nz1))
    g[:, :jee] = planes(cfg, jee, 6)
    ref = np.zeros_like(g)
    ob.op("differzy", ptr(g), ptr(ref), which, cfg.nx)
    geni = np.ascontiguousarray(g[3, :jee])
    res = np.zeros((cfg.ny + 1, cfg.nz1))
    dummy = np.zeros(16)
    L.differzy_(p(geni), p(res), p(dummy), p(dummy), p(dummy), iref(chv), iref(1), iref(f1), iref(f2), iref(jee))
    assert np.abs(res[:jee] - ref[3, :jee]).max() < OP_TOL * np.abs(ref[3, :jee]).max()


@pytest.mark.parametrize("uu", [True, False], ids=["u_rows", "v_rows"])
def test_vistyy_and_implzy_dropins(fbox, uu):
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    jee = cfg.ny + 1 if uu else cfg.ny
    u, r = planes(cfg, jee, 7), planes(cfg, jee, 8)
    ref = np.zeros_like(u)
    ob.op("vistyy", ptr(u), ptr(ref), jee, cfg.nx)
    res = np.zeros((cfg.ny + 1, cfg.nz1))
    dummy = np.zeros(16)
    u2 = np.ascontiguousarray(u[5])
    L.vistyy_(p(u2), p(res), p(dummy), p(dummy), p(dummy), iref(jee))
    assert np.abs(res[:jee] - ref[5]).max() < OP_TOL * np.abs(ref[5]).max()
    rkk = -0.02
    ui = u.copy()
    ob.op("implzy", ptr(ui), ptr(r), jee, rkk, cfg.nx)
    v = np.ascontiguousarray(u[5])
    r5 = np.ascontiguousarray(r[5])
    L.implzy_(p(v), p(r5), p(dummy), p(dummy), iref(jee), dref(rkk))
    assert np.abs(v - ui[5]).max() < OP_TOL * np.abs(ui[5]).max()


def test_fourxz_dropin(fbox, oracle_lib):
    """Rows jbf..jef of complex (0:nz2) spectra and (0:nz/2) physical rows,
    both directions, against the oracle's c2r / r2c (ffts.f90:27-67)."""
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    nr = 5
    rng = np.random.default_rng(9)
    fou = np.ascontiguousarray(rng.standard_normal((nr, cfg.nz1)))
    phys = np.zeros((nr, cfg.nz + 2))
    L.fourxz_(p(fou), p(phys), iref(1), iref(cfg.ny + 1), iref(3), iref(3 + nr - 1))
    for r in range(nr):
        want = np.zeros(cfg.nz)
        oracle_lib.lib().oc_c2r(ob.h, ptr(np.ascontiguousarray(fou[r])), ptr(want))
        assert np.abs(phys[r, :cfg.nz] - want).max() < OP_TOL * np.abs(want).max()
    back = np.zeros_like(fou)
    L.fourxz_(p(back), p(phys), iref(-1), iref(cfg.ny + 1), iref(3), iref(3 + nr - 1))
    for r in range(nr):
        want = np.zeros(cfg.nz1)
        oracle_lib.lib().oc_r2c(ob.h, ptr(np.ascontiguousarray(phys[r, :cfg.nz])), ptr(want))
        assert np.abs(back[r] - want).max() < OP_TOL * np.abs(want).max()


@pytest.mark.parametrize("iopt", [-1, 1])
def test_cosftx_dropin(fbox, oracle_lib, iopt):
    """cosftx(res(2, mpvb), nx, mpvb, mpve, iopt) as pois.f90:90 calls it:
    lines of stride nx starting at the second point."""
    L, cfg, st, ob = fbox
    from oracle.oracle import ptr
    nl = 7
    c = np.ascontiguousarray(np.random.default_rng(10).standard_normal((nl, cfg.nx)))
    want = np.ascontiguousarray(c[:, 1:])
    oracle_lib.lib().oc_cosft(cfg.nx - 1, nl, ptr(want), iopt)
    L.cosfti_(iref(cfg.nx - 1))
    sub = c.ctypes.data + 8        # c(2, mpvb)
    L.cosftx_(C.c_void_p(sub), iref(cfg.nx), iref(1), iref(nl), iref(iopt))
    assert np.abs(c[:, 1:] - want).max() < OP_TOL * np.abs(want).max()


def test_soltriy_dropin(fbox):
    """soltriy(u, us, ii, a) (pois.f90:211-239): column ii of a complex
    plane solved with the tridiagonal a, which comes back LU-factored; in
    place as pois.f90:116 calls it.  Against a dense solve."""
    L, cfg, st, ob = fbox
    ny1 = cfg.ny - 1
    rng = np.random.default_rng(11)
    a = np.zeros((ny1, 3))
    a[:, 0] = rng.uniform(0.5, 1.0, ny1)
    a[:, 2] = rng.uniform(0.5, 1.0, ny1)
    a[:, 1] = -(a[:, 0] + a[:, 2] + 1.0)
    A = np.diag(a[:, 1]) + np.diag(a[1:, 0], -1) + np.diag(a[:-1, 2], 1)
    u = np.ascontiguousarray(rng.standard_normal((ny1, cfg.nz1)))
    u0 = u.copy()
    ii = 2
    af = a.copy()  # a stays the unfactored input
    L.soltriy_(p(u), p(u), iref(ii), p(af))
    x = np.linalg.solve(A, u0[:, 2 * ii] + 1j * u0[:, 2 * ii + 1])
    assert np.abs(u[:, 2 * ii] + 1j * u[:, 2 * ii + 1] - x).max() < 1e-12 * np.abs(x).max()
    others = [k for k in range(cfg.nz1) if k not in (2 * ii, 2 * ii + 1)]
    assert np.array_equal(u[:, others], u0[:, others])
    # the factored diagonal holds the inverse pivots
    assert af[-1, 1] != a[-1, 1]


def test_coefinv_and_mtr_dropins(fbox):
    """coefinv of the unfactored implicit matrix vyui (the copy made before
    coeft.f90:1131-1134 factors viyu) equals the box's factored viyu; mtr
    returns the box's ayp (coeft.f90:1244-1332) for the box's grid."""
    L, cfg, st, ob = fbox
    vyui = np.ascontiguousarray(ob.coef("vyui").T)            # (n, 3): Fortran a(3, n)
    L.coefinv_(p(vyui), iref(cfg.ny + 1), iref(1))
    assert np.abs(vyui.T - ob.coef("viyu")).max() < 1e-14 * np.abs(ob.coef("viyu")).max()
    ayp = np.zeros((cfg.ny, 3))
    y = np.ascontiguousarray(st["y"], np.float64)
    L.mtr_(p(ayp), iref(1), iref(1), p(y))
    want = ob.coef("ayp")
    assert np.abs(ayp[:cfg.ny - 1].T - want).max() < 1e-14 * np.abs(want).max()


def test_chp2x_chx2p_dropins(fbox):
    """Planes to pencils and back (change.f90:224-275), one process."""
    L, cfg, st, ob = fbox
    R = cfg.ny
    g = planes(cfg, R, 12)
    pen = np.zeros((cfg.nz1 * R, cfg.nx))
    buf = np.zeros(1)
    L.chp2x_(p(pen), p(g), p(buf), iref(0), iref(R), iref(0))
    assert np.array_equal(pen, to_pencils(g))
    back = pen.copy()
    L.chx2p_(p(back), p(pen), p(buf), iref(0), iref(R), iref(0))
    assert np.array_equal(back.reshape(-1), g.reshape(-1))


def test_statsp_dropin(fbox, oracle_lib):
    """statsp (stotf.F90:51-60) called with the reference's argument list:
    one sample of the box's state, against the oracle's statsp on the same
    state (the box's fields are set through the C ABI)."""
    from opentbl_amd.lib import FIELDS, NSTAT, lib
    L, cfg, st, _ = fbox
    L.otbl_f_box.restype = C.c_void_p
    box = C.c_void_p(L.otbl_f_box())
    ob = oracle_lib.OracleBox(cfg, st)
    for n in ("u", "v", "w", "p"):
        a = np.ascontiguousarray(st[n], np.complex128).view(np.float64)
        assert lib().otbl_set_field(box, FIELDS[n], a.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert lib().otbl_stats_enable(box, 1) == 0
    dummy = np.zeros(8)
    args = [p(dummy)] * 21 + [iref(0), iref(0)]
    L.statsp_(*args)
    L.statsp_(*args)
    ob.statsp()
    ob.statsp()
    want = ob.stats()
    ical = C.c_int()
    assert lib().otbl_stats_ical(box, C.byref(ical)) == 0 and ical.value == 2
    for q in range(NSTAT):
        name = lib().otbl_stats_name(q).decode()
        g = np.zeros(cfg.nx * (cfg.ny + 1))
        assert lib().otbl_stats_get(box, q, g.ctypes.data_as(C.POINTER(C.c_double))) == 0
        w = want[name]
        assert np.abs(g.reshape(w.shape) - w).max() <= 1e-10 * max(np.abs(w).max(), 1e-300), name
    assert lib().otbl_stats_enable(box, 0) == 0


@pytest.mark.parametrize("inplace", [True, False])
def test_soltrix_dropin(fbox, inplace):
    """soltrix(u, nstr, nlen, neq, us, npter1, npter2, a, iflag)
    (coeft.f90:1926-1967): columns npter2.. of u(nstr, *), rows npter1.., a
    dense solve each; a comes back as coefinv factors it and iflag = 1; the
    rows of us outside the system are untouched.  Then a second call with
    iflag = 1 reuses the factors.  (The reference calls it in place with
    one column, coeft.f90:1208,1225.)"""
    L, cfg, st, ob = fbox
    rng = np.random.default_rng(17)
    nstr, ncols, neq, npter1, npter2, nlen = 30, 6, 24, 4, 2, 4
    a = np.zeros((nstr, 3))                       # Fortran a(3, nstr) as C (nstr, 3)
    a[:, 0] = rng.uniform(0.3, 1.0, nstr)
    a[:, 2] = rng.uniform(0.3, 1.0, nstr)
    a[:, 1] = -(a[:, 0] + a[:, 2] + 0.5)
    r = slice(npter1 - 1, npter1 - 1 + neq)
    A = np.diag(a[r, 1]) + np.diag(a[r, 0][1:], -1) + np.diag(a[r, 2][:-1], 1)
    u = np.ascontiguousarray(rng.standard_normal((ncols, nstr)))   # Fortran u(nstr, ncols)
    us = u if inplace else np.ascontiguousarray(rng.standard_normal((ncols, nstr)))
    u0, us0 = u.copy(), us.copy()
    af = a.copy()
    fl = C.c_int(0)
    L.soltrix_(p(u), iref(nstr), iref(nlen), iref(neq), p(us), iref(npter1), iref(npter2), p(af), C.byref(fl))
    assert fl.value == 1
    for j in range(ncols):
        if npter2 - 1 <= j < npter2 - 1 + nlen:
            x = np.linalg.solve(A, u0[j, r])
            assert np.abs(us[j, r] - x).max() <= 1e-12 * np.abs(x).max(), j
            outside = np.r_[0:npter1 - 1, npter1 - 1 + neq:nstr]
            assert np.array_equal(us[j, outside], us0[j, outside]), j
        else:
            assert np.array_equal(us[j], us0[j]), j
    # the factors are coefinv's (the same loop, coeft.f90:1340-1352)
    ref = np.ascontiguousarray(a[r].copy())
    L.coefinv_(p(ref), iref(neq), iref(1))
    assert np.array_equal(af[r], ref)
    # iflag = 1: the factored a is used as it is
    v = np.ascontiguousarray(rng.standard_normal((ncols, nstr)))
    v0 = v.copy()
    L.soltrix_(p(v), iref(nstr), iref(1), iref(neq), p(v), iref(npter1), iref(1), p(af), C.byref(fl))
    x = np.linalg.solve(A, v0[0, r])
    assert np.abs(v[0, r] - x).max() <= 1e-12 * np.abs(x).max()


def test_coefb_and_bookkeeping_dropins(fbox, oracle_lib):
    """coefb (coeft.f90:1360-1789) gives the oracle's closure weights for
    every flag the reference uses; pointersfft, pointers_p2p and comm_setup
    (OpenMP / MPI bookkeeping, ffts.f90:188, change.f90:29,150) are
    accepted and leave their arguments alone."""
    L, cfg, st, ob = fbox
    O = oracle_lib.lib()
    for flag, n in {0: 3, 1: 5, 3: 4, 4: 4, 5: 4, 6: 4, 7: 4}.items():
        sp = [0.7, 0.9, 1.1, 1.3]
        got, want = np.zeros(5), np.zeros(5)
        L.coefb_(iref(n), iref(5), dref(sp[0]), dref(sp[1]), dref(sp[2]), dref(sp[3]), p(got), iref(flag))
        O.oc_coefb(n, flag, *sp, p(want))
        assert np.abs(got[:n] - want[:n]).max() <= 1e-12 * np.abs(want[:n]).max(), flag
    rank = C.c_int(0)
    L.pointersfft_(C.byref(rank))
    L.pointers_p2p_(C.byref(rank))
    node, domain = np.zeros(8, np.int32), np.zeros(8, np.int32)
    L.comm_setup_(iref(cfg.nx), iref(cfg.ny), iref(cfg.nz1), C.byref(rank), iref(1),
                  node.ctypes.data_as(C.POINTER(C.c_int)), domain.ctypes.data_as(C.POINTER(C.c_int)), iref(1))
    assert rank.value == 0 and not node.any() and not domain.any()
