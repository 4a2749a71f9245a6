"""ctypes view of the C oracle (liboracle_impactz.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
Particle arrays are (N, 9) float64 C-contiguous numpy arrays, i.e. the
reference's Pts1(9, N) memory layout; grids are numpy arrays of shape
(Nz, Ny, Nx) (x fastest, the reference's rho(Nx, Ny, Nz) memory layout).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_impactz.so")
SRC_PATH = os.path.join(HERE, "impactz_oracle.c")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class OrConst(C.Structure):
    _fields_ = [("clight", C.c_double), ("pi", C.c_double),
                ("scxl", C.c_double), ("scfreq", C.c_double)]


NVAL = 14


class OrRFData(C.Structure):
    _fields_ = [("fourier", C.c_int), ("n", C.c_int), ("zdat", _dp), ("edat", _dp),
                ("epdat", _dp), ("fcoef", _dp)]


class OrElement(C.Structure):
    _fields_ = [("bitype", C.c_int), ("mapstp", C.c_int), ("v", C.c_double * NVAL),
                ("length", C.c_double), ("zedge", C.c_double),
                ("rf", C.POINTER(OrRFData))]


class RFData:
    """The Data module's arrays for one rfdata file: read1_Data's table with
    z shifted to start at 0 (src/DataStruct/Data.f90:183-199) or read3_Data's
    coefficients padded to an odd count (:387-406)."""

    def __init__(self, field, fourier: bool):
        a = np.asarray(field, dtype=np.float64)
        self.c = OrRFData()
        self.c.fourier = int(fourier)
        if fourier:
            coef = a.ravel().copy()
            if coef.size % 2 == 0:
                coef = np.append(coef, 0.0)
            self.keep = [coef]
            self.c.n = coef.size
            self.c.fcoef = _p(coef)
        else:
            t = a.reshape(-1, 4)
            cols = [np.ascontiguousarray(t[:, k]) for k in range(4)]
            cols[0] = cols[0] - t[0, 0]
            self.keep = cols
            self.c.n = t.shape[0]
            self.c.zdat, self.c.edat, self.c.epdat = (_p(cols[k]) for k in range(3))


class OrWake(C.Structure):
    _fields_ = [("kind", C.c_int), ("aa", C.c_double), ("gg", C.c_double),
                ("leng", C.c_double), ("n", C.c_int), ("zdat", _dp), ("edat", _dp),
                ("epdat", _dp), ("eppdat", _dp), ("flagbtw", C.c_int),
                ("scale", C.c_double), ("r0", C.c_double), ("zoff", C.c_double),
                ("blength", C.c_double)]


def make_wake(kind, aa=0.0, gg=0.0, leng=0.0, table=None, flagbtw=0, scale=1.0,
              r0=0.0, zoff=0.0, blength=0.0):
    """An or_wake: kind 1 analytic (aa, gg, leng), kind 2 the rows of a
    read1wk_Data file (z shifted to start at 0, src/DataStruct/Data.f90:
    613-617), kind 3 the CSR wake of a bend (r0, zoff, blength)."""
    w = OrWake()
    w.kind, w.aa, w.gg, w.leng, w.flagbtw, w.scale = kind, aa, gg, leng, flagbtw, scale
    w.r0, w.zoff, w.blength = r0, zoff, blength
    if table is not None:
        t = np.asarray(table, dtype=np.float64).reshape(-1, 4)
        cols = [np.ascontiguousarray(t[:, k]) for k in range(4)]
        cols[0] = cols[0] - t[0, 0]
        w._keep = cols
        w.n = t.shape[0]
        w.zdat, w.edat, w.epdat, w.eppdat = (_p(x) for x in cols)
    return w


def wakefield(c, pts, rng, nz, current, npt, gam, charge, w: OrWake):
    ex, ey, ez = np.zeros(nz), np.zeros(nz), np.zeros(nz)
    p = np.ascontiguousarray(pts, dtype=np.float64)
    if lib().or_wakefield(C.byref(c), _p(p), p.shape[0],
                          _p(np.ascontiguousarray(rng, dtype=np.float64)), nz, current,
                          npt, gam, charge, C.byref(w), _p(ex), _p(ey), _p(ez)) != 0:
        raise ValueError("bad wake kind")
    return np.stack([ex, ey, ez])


def csr_wake(r0, ptmin, hx, blength, rho, gam):
    """csrwakeTrIGF_FieldQuant: the CSR ezwake (V/m) of line density rho."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    ez = np.zeros(rho.shape[0])
    lib().or_csr_wake(rho.shape[0], r0, ptmin, hx, blength, _p(rho), gam, _p(ez))
    return ez


def kick1wake(c, pts, phi, rng, ms, refpt, curr, tau, mass, chrg, wake):
    """kick1wake_BeamBunch in place on pts; wake = (exwake, eywake, ezwake)."""
    nz, ny, nx = phi.shape
    w = np.ascontiguousarray(wake, dtype=np.float64)
    lib().or_kick1wake(C.byref(c), _p(pts), pts.shape[0], nx, ny, nz, _p(phi),
                       _p(np.ascontiguousarray(rng, dtype=np.float64)),
                       _p(np.ascontiguousarray(ms, dtype=np.float64)),
                       _p(np.ascontiguousarray(refpt, dtype=np.float64)), curr, tau, mass,
                       chrg, _p(w[0].copy()), _p(w[1].copy()), _p(w[2].copy()))


def make_element(bitype, v, mapstp=1, length=0.0, zedge=0.0, rf: RFData | None = None):
    """An or_element; v = the element's values (a lone number is a
    quadrupole gradient)."""
    e = OrElement()
    e.bitype, e.mapstp, e.length, e.zedge = bitype, mapstp, length, zedge
    for k, x in enumerate(element_values(v)):
        e.v[k] = x
    if rf is not None:
        e.rf = C.pointer(rf.c)
    return e


class OrState(C.Structure):
    _fields_ = [("bcurr", C.c_double), ("current", C.c_double),
                ("mass", C.c_double), ("charge", C.c_double),
                ("npt", C.c_int64), ("refpt", C.c_double * 6),
                ("z", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction, like the reference's x86 build)."""
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(SRC_PATH):
        subprocess.check_call(["gcc", "-O2", "-std=c99", "-ffp-contract=off",
                               "-fPIC", "-shared", SRC_PATH, "-o", LIB_PATH,
                               "-lm"])
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_physconst.argtypes = [C.c_double, C.POINTER(OrConst)]
        L.or_quad_map.argtypes = [C.POINTER(OrConst), C.c_double, C.c_double,
                                  C.c_int, _dp, C.c_double, C.c_double, _dp]
        L.or_element_map.argtypes = [C.POINTER(OrConst), C.c_double, C.c_double,
                                     C.POINTER(OrElement), _dp, C.c_double, C.c_double,
                                     _dp]
        L.or_element_map.restype = C.c_int
        L.or_map1.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp, _dp, C.c_double,
                              C.POINTER(OrElement), C.c_double, C.c_double]
        L.or_map1.restype = C.c_int
        L.or_dipole_face.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp,
                                     C.POINTER(OrElement), C.c_int, C.c_double, C.c_double]
        L.or_sector_dipole.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp, _dp,
                                       C.c_double, C.POINTER(OrElement), C.c_double,
                                       C.c_double]
        L.or_isr_chicane.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp, C.c_double,
                                     C.POINTER(OrElement), C.c_double, _dp]
        L.or_conv1st.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp,
                                 C.c_double, C.c_double, _dp, _i64p, _dp]
        L.or_conv1st.restype = C.c_int64
        L.or_update_compdom.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_deposit.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_int,
                                 _dp, _dp, _dp]
        L.or_green_igf.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, _dp]
        L.or_solve_openbc.argtypes = [_dp, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, _dp]
        L.or_solve_openbc.restype = C.c_int
        L.or_solve_direct.argtypes = [_dp, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, _dp]
        L.or_cvbkforth1st.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp]
        L.or_efield.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.or_kick1.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, C.c_int,
                               C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                               C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_diag_sums.argtypes = [_dp, C.c_int64, _dp, _dp, _dp]
        L.or_chgupdate.argtypes = [_dp, C.c_int64, C.c_int, _dp,
                                   C.POINTER(C.c_int64)]
        L.or_diag_row.argtypes = [C.POINTER(OrConst), _dp, _dp, C.c_double,
                                  C.c_int64, _dp, C.c_double, _dp]
        L.or_step.argtypes = [C.POINTER(OrConst), _dp, C.c_int64,
                              C.POINTER(OrState), C.c_int, C.c_int, C.c_int,
                              C.c_double, C.POINTER(OrElement), C.c_double,
                              C.POINTER(OrWake), _dp, _dp]
        L.or_wakefield.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp, C.c_int,
                                   C.c_double, C.c_int64, C.c_double, C.c_double,
                                   C.POINTER(OrWake), _dp, _dp, _dp]
        L.or_wakefield.restype = C.c_int
        L.or_kick1wake.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, C.c_int, C.c_int,
                                   C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_step.restype = C.c_int64
        L.or_drift2.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, _dp, _dp, C.c_double]
        L.or_elem_field.argtypes = [C.POINTER(OrConst), C.c_int, _dp, C.c_double,
                                    C.c_double, C.POINTER(OrRFData), _dp, _dp]
        L.or_elem_field.restype = C.c_int
        L.or_kick2.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, C.c_int, C.c_int,
                               C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_int, _dp, C.c_double,
                               C.c_double, C.POINTER(OrRFData), C.c_double, _dp]
        L.or_kick2.restype = C.c_int
        L.or_step2.argtypes = [C.POINTER(OrConst), _dp, C.c_int64, C.POINTER(OrState),
                               C.c_int, C.c_int, C.c_int, C.c_double,
                               C.POINTER(OrElement), C.c_double, C.POINTER(OrWake), _dp, _dp]
        L.or_step2.restype = C.c_int64
        L.or_csr_wake.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                  _dp, C.c_double, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def physconst(freq: float) -> OrConst:
    c = OrConst()
    lib().or_physconst(freq, C.byref(c))
    return c


def quad_map(c: OrConst, tau, grad, mapstp, refpt, charge, mass):
    refpt = np.ascontiguousarray(refpt, dtype=np.float64).copy()
    xm = np.zeros(36)
    lib().or_quad_map(C.byref(c), tau, grad, mapstp, _p(refpt), charge, mass, _p(xm))
    return xm.reshape(6, 6), refpt


def element_values(v) -> np.ndarray:
    """v1..v14 of an element, zero-padded (a lone number is a quadrupole
    gradient)."""
    out = np.zeros(NVAL)
    if np.isscalar(v):
        out[0] = float(v)
        return out
    vals = np.asarray(v, dtype=np.float64).ravel()[:NVAL]
    out[: vals.size] = vals
    return out


def element_map(c, tau, bitype, v, mapstp, refpt, charge, mass, t=0.0, length=0.0,
                zedge=0.0, rf: RFData | None = None):
    refpt = np.ascontiguousarray(refpt, dtype=np.float64).copy()
    xm = np.zeros(36)
    e = make_element(bitype, v, mapstp, length, zedge, rf)
    if lib().or_element_map(C.byref(c), t, tau, C.byref(e), _p(refpt), charge, mass,
                            _p(xm)) != 0:
        raise ValueError(f"no linear map for element type {bitype}")
    return xm.reshape(6, 6), refpt


def map1(c, pts, refpt, z, tau, bitype, v, mapstp, charge, mass, length=0.0, zedge=0.0,
         rf: RFData | None = None):
    """drift1_BeamBunch in place on pts; v = the element's values (or a
    quadrupole gradient).  Returns the new z."""
    zz = C.c_double(z)
    e = make_element(bitype, v, mapstp, length, zedge, rf)
    if lib().or_map1(C.byref(c), _p(pts), pts.shape[0], _p(refpt), C.byref(zz),
                     tau, C.byref(e), charge, mass) != 0:
        raise ValueError(f"element type {bitype} with values {list(e.v)} is not restated")
    return zz.value


def dipole_face(c, pts, refpt, length, v, exit, charge, mass):
    """Fpol_Dipole (exit=False, with the bend's refpt(5) advance) or
    Bpol_Dipole (exit=True) in place."""
    e = make_element(4, v, 1, length)
    lib().or_dipole_face(C.byref(c), _p(pts), pts.shape[0], _p(refpt), C.byref(e),
                         int(bool(exit)), charge, mass)


def sector_dipole(c, pts, refpt, z, tau, length, v, charge, mass):
    """Sector_Dipole over tau in place; returns the new z."""
    zz = C.c_double(z)
    e = make_element(4, v, 1, length)
    lib().or_sector_dipole(C.byref(c), _p(pts), pts.shape[0], _p(refpt), C.byref(zz),
                           tau, C.byref(e), charge, mass)
    return zz.value


def isr_chicane(c, pts, refpt, tau, length, v, mass, rd):
    """ISRchicane over tau in place with deviates rd (one per row of pts)."""
    e = make_element(4, v, 1, length)
    rd = np.ascontiguousarray(rd, dtype=np.float64)
    assert rd.shape == (pts.shape[0],)
    lib().or_isr_chicane(C.byref(c), _p(pts), pts.shape[0], _p(refpt), tau, C.byref(e),
                         mass, _p(rd))


def drift2(c, pts, refpt, z, tau):
    """drift2_BeamBunch in place; returns the new z."""
    zz = C.c_double(z)
    lib().or_drift2(C.byref(c), _p(pts), pts.shape[0], _p(refpt), C.byref(zz), tau)
    return zz.value


def elem_field(c, bitype, v, length, zedge, pos, rf: RFData | None = None):
    f = np.zeros(6)
    vv = element_values(v)
    p = np.ascontiguousarray(pos, dtype=np.float64)
    if lib().or_elem_field(C.byref(c), bitype, _p(vv), length, zedge,
                           C.byref(rf.c) if rf is not None else None, _p(p), _p(f)) != 0:
        raise ValueError(f"no analytic field for element type {bitype}")
    return f


def kick2(c, pts, phi, rng, ms, refpt, curr, tau, mass, chrg, bitype, v, length, zedge, z,
          rf: RFData | None = None, wake=None):
    """kick2_BeamBunch in place on pts (and refpt(6)); with wake =
    (exwake, eywake, ezwake) kick2wake_BeamBunch."""
    wk = None if wake is None else np.ascontiguousarray(wake, dtype=np.float64).ravel()
    nz, ny, nx = phi.shape
    vv = element_values(v)
    if lib().or_kick2(C.byref(c), _p(pts), pts.shape[0], nx, ny, nz, _p(phi),
                      _p(np.ascontiguousarray(rng, dtype=np.float64)),
                      _p(np.ascontiguousarray(ms, dtype=np.float64)), _p(refpt), curr, tau,
                      mass, chrg, bitype, _p(vv), length, zedge,
                      C.byref(rf.c) if rf is not None else None, z,
                      _p(wk) if wk is not None else None) != 0:
        raise ValueError(f"no analytic field for element type {bitype}")


def conv1st(c, pts, refpt, xrad, yrad, npt, current):
    """In place on pts; returns (n_after, ptrange, npt, current)."""
    pr = np.zeros(6)
    npt_ = C.c_int64(npt)
    cur = C.c_double(current)
    n = lib().or_conv1st(C.byref(c), _p(pts), pts.shape[0], _p(refpt), xrad,
                         yrad, _p(pr), C.byref(npt_), C.byref(cur))
    return n, pr, npt_.value, cur.value


def update_compdom(ptrange, grid):
    rng, ms = np.zeros(6), np.zeros(3)
    lib().or_update_compdom(_p(np.ascontiguousarray(ptrange, dtype=np.float64)),
                            grid[0], grid[1], grid[2], _p(rng), _p(ms))
    return rng, ms


def deposit(pts, grid, rng, ms):
    nx, ny, nz = grid
    rho = np.zeros((nz, ny, nx))
    lib().or_deposit(_p(pts), pts.shape[0], nx, ny, nz, _p(rng), _p(ms), _p(rho))
    return rho


def green_igf(grid, h):
    nx, ny, nz = grid
    g = np.zeros((nz + 1, ny + 1, nx + 1))
    lib().or_green_igf(nx, ny, nz, h[0], h[1], h[2], _p(g))
    return g


def solve_openbc(rho, h):
    nz, ny, nx = rho.shape
    phi = np.zeros_like(rho)
    rc = lib().or_solve_openbc(_p(np.ascontiguousarray(rho)), nx, ny, nz,
                               h[0], h[1], h[2], _p(phi))
    if rc != 0:
        raise ValueError(f"or_solve_openbc failed ({rc})")
    return phi


def solve_direct(rho, h):
    nz, ny, nx = rho.shape
    phi = np.zeros_like(rho)
    lib().or_solve_direct(_p(np.ascontiguousarray(rho)), nx, ny, nz,
                          h[0], h[1], h[2], _p(phi))
    return phi


def cvbkforth1st(c, pts, refpt):
    lib().or_cvbkforth1st(C.byref(c), _p(pts), pts.shape[0], _p(refpt))


def efield(phi, ms):
    nz, ny, nx = phi.shape
    e = np.zeros((3,) + phi.shape)
    lib().or_efield(_p(np.ascontiguousarray(phi)), nx, ny, nz, _p(ms),
                    _p(e[0]), _p(e[1]), _p(e[2]))
    return e


def kick1(c, pts, phi, rng, ms, refpt, curr, tau, mass, chrg):
    nz, ny, nx = phi.shape
    lib().or_kick1(C.byref(c), _p(pts), pts.shape[0], nx, ny, nz,
                   _p(np.ascontiguousarray(phi)), _p(rng), _p(ms), _p(refpt),
                   curr, tau, mass, chrg)


def diag_sums(pts):
    s, m, r2 = np.zeros(37), np.zeros(6), C.c_double(0)
    lib().or_diag_sums(_p(pts), pts.shape[0], _p(s), _p(m), C.byref(r2))
    return s, m, r2.value


def chgupdate(pts, qmcclist):
    """chgupdate_BeamBunch: particles of each charge state."""
    q = np.ascontiguousarray(qmcclist, dtype=np.float64)
    out = (C.c_int64 * len(q))()
    p = np.ascontiguousarray(pts, dtype=np.float64)
    lib().or_chgupdate(_p(p), p.shape[0], len(q), _p(q), out)
    return np.array(out[:], dtype=np.int64)


def diag_row(c, sums, amax, r2max, npt, refpt, mass):
    out = np.zeros(41)
    lib().or_diag_row(C.byref(c), _p(np.ascontiguousarray(sums)),
                      _p(np.ascontiguousarray(amax)), r2max, npt, _p(refpt),
                      mass, _p(out))
    return out


_SC_CURRENT = float(np.float32(1.0e-10))   # the loop's REAL(4) space-charge threshold


class Run:
    """The reference segment loop over a lattice, on the C oracle."""

    def __init__(self, pts, beam, lattice, grid, flagmap: int = 1):
        self.flagmap = flagmap
        self.zedge = 0.0
        self._left = 0   # segments left in the current element
        self._rf = {}
        # the loop's wake switches: flagwake, flagwakeread, flagbtw, scwk,
        # aawk/ggwk/lengwk and the read-in table (src/Contrl/AccSimulator.f90:656-720)
        self.wk = {"on": 0, "read": 0, "btw": 0, "scale": 1.0, "aa": 0.0, "gg": 0.0,
                   "leng": 0.0, "table": None}
        # the loop's CSR bookkeeping (src/Contrl/AccSimulator.f90:722-740,
        # 1032-1056, 1131-1138, 1374-1385, 1498-1499): zbleng, sentr,
        # flagbendcsr, bendlencsr and the last bend's radius r0
        self.csr = {"zbleng": 0.0, "sentr": 0.0, "flagbend": 0, "bendlen": 0.0, "r0": 0.0}
        # ISRchicane's standard-normal deviates (the reference draws them from
        # the compiler's random_number): a generator the caller seeds, one
        # deviate per particle id per bend step
        self.isr_rng = None
        self.c = physconst(beam.freq)
        self.pts = np.ascontiguousarray(pts, dtype=np.float64).copy()
        self.n = self.n0 = self.pts.shape[0]
        self.grid = tuple(grid)
        self.lattice = lattice
        st = OrState()
        st.bcurr = beam.current
        st.current = beam.current
        st.mass = beam.mass
        st.charge = beam.charge
        st.npt = self.n
        for k in range(6):
            st.refpt[k] = 0.0
        st.refpt[4] = beam.phase0
        st.refpt[5] = -(beam.kinetic / beam.mass + 1.0)
        st.z = 0.0
        self.st = st

    def segments(self):
        """The lattice's steps; a -41 element switches the read-in wake as
        the loop reaches it."""
        for e in self.lattice:
            if e.type == -41:
                self.wake_switch(e)
                continue
            for _ in range(e.nseg):
                yield e

    def wake_switch(self, e):
        """A -41 element (src/Contrl/AccSimulator.f90:960-980)."""
        if self.wk["read"] == 0:
            self.wk["on"] = 0   # :759, for every element
        v = e.values
        if v[1] > 0.0:
            self.wk["table"] = e.field
        if v[2] > 0.0:
            self.wk.update(on=1, read=1, scale=v[0])
            if v[2] < 10.0:
                self.wk["btw"] = 2
        else:
            self.wk.update(on=0, read=0, scale=1.0)

    def _wake_entry(self, e):
        """The wake switches on entering an element (src/Contrl/
        AccSimulator.f90:759,859-875)."""
        if self.wk["read"] == 0:
            self.wk["on"] = 0
        if e.type in (105, 106):
            v = e.values
            self.wk["aa"], self.wk["gg"], self.wk["leng"] = v[11], v[12], v[13]
            self.wk["on"] = 1 if v[13] > 0.0 else 0

    def _csr_entry(self, e):
        """Entering a bend (src/Contrl/AccSimulator.f90:1032-1056,1066-1091):
        the transient-drift flag, the bend's length for the elements after
        it, sentr reset, and r0 = |1/hd0|."""
        if e.type != 4:
            return
        v = e.values
        if self.st.bcurr > 0.0 and v[2] > 500:
            self.csr.update(flagbend=1, bendlen=e.length, sentr=0.0)
        else:
            self.csr["flagbend"] = 0
        hd0 = v[0] / e.length
        self.csr["r0"] = abs(1.0 / hd0)

    def _csr_exit(self, e):
        """After an element's last step (:1498-1499)."""
        self.csr["zbleng"] += e.length
        self.csr["sentr"] += e.length

    def csr_wake(self, e, z):
        """The CSR or_wake of a step of e whose first half push left z (the
        loop's flagcsr, zwkmin and bendlen, :1131-1138,1374-1385), or None."""
        k = self.csr
        if e.type == 4:
            if not e.values[2] > 200:
                return None
            return make_wake(3, r0=k["r0"], zoff=z - k["zbleng"], blength=e.length)
        if not k["flagbend"]:
            return None
        return make_wake(3, r0=k["r0"], zoff=z - k["zbleng"] + k["sentr"],
                         blength=k["bendlen"])

    def wake(self):
        """The or_wake of the current step, or None."""
        w = self.wk
        if not w["on"]:
            return None
        if w["read"]:
            return make_wake(2, table=w["table"], flagbtw=w["btw"], scale=w["scale"])
        return make_wake(1, w["aa"], w["gg"], w["leng"], scale=w["scale"])

    def rfdata(self, e):
        """The element's field arrays (cached per element)."""
        if e.field is None:
            return None
        # read3_Data's coefficients for SolRF, and for every cavity under the
        # Lorentz integrator (src/Contrl/AccSimulator.f90:826-847)
        fourier = e.type == 105 or self.flagmap != 1
        key = (id(e.field), fourier)
        if key not in self._rf:
            self._rf[key] = (RFData(e.field, fourier=fourier), e.field)
        return self._rf[key][0]

    def step(self, e, want_fields=False):
        nx, ny, nz = self.grid
        rho = np.zeros((nz, ny, nx)) if want_fields else None
        phi = np.zeros((nz, ny, nx)) if want_fields else None
        self.finish_element()
        tau1 = 0.5 * e.length / e.nseg
        if self._left == 0:
            # a new element: its entrance z (src/Contrl/AccSimulator.f90:1025)
            self._left, self.zedge = e.nseg, self.st.z
            self._wake_entry(e)
            self._csr_entry(e)
        self._left -= 1
        if e.type == 4:
            out = self._bend_step(e, tau1, entering=self._left == e.nseg - 1, rho=rho,
                                  phi=phi)
            if self._left == 0:
                self._csr_exit(e)
            return out
        el = make_element(e.type, e.values, e.mapstp, e.length, self.zedge, self.rfdata(e))
        out = (_p(rho) if want_fields else None, _p(phi) if want_fields else None)
        if self.flagmap == 1:
            # map1 moves z by tau1 before the field solve the wake follows
            w = self.csr_wake(e, self.st.z + tau1) if self.st.bcurr >= _SC_CURRENT else None
            w = w if w is not None else self.wake()
            n = lib().or_step(C.byref(self.c), _p(self.pts), self.n, C.byref(self.st), nx,
                              ny, nz, tau1, C.byref(el), e.radius,
                              C.byref(w) if w is not None else None, *out)
        else:
            w = self.wake()
            n = lib().or_step2(C.byref(self.c), _p(self.pts), self.n, C.byref(self.st), nx,
                               ny, nz, tau1, C.byref(el), e.radius,
                               C.byref(w) if w is not None else None, *out)
        if n < 0:
            raise RuntimeError("or_step failed")
        self.n = n
        if self._left == 0:
            self._csr_exit(e)
        return rho, phi

    def _bend_step(self, e, tau1, entering, rho=None, phi=None):
        """A bend segment (src/Contrl/AccSimulator.f90:1032-1091, 1114-1138,
        1141-1396, 1417-1448, 1491-1494): the entrance face first, a half
        sector map, with current the space-charge block (pipe radius 100 m,
        the CSR wake when v3 > 200) and kick, ISRchicane with the caller's
        deviates, the second half sector map, the exit face last.  Zero
        current runs under both integrators; current under Flagmap=1."""
        st = self.st
        pts = self.pts[: self.n]
        refpt = np.ctypeslib.as_array(st.refpt)
        if entering:
            dipole_face(self.c, pts, refpt, e.length, e.values, False, st.charge, st.mass)
        z = sector_dipole(self.c, pts, refpt, st.z, tau1, e.length, e.values, st.charge,
                          st.mass)
        if st.bcurr >= _SC_CURRENT:
            if self.flagmap != 1:
                raise ValueError("bends with current are restated under Flagmap=1")
            n, pr, st.npt, st.current = conv1st(self.c, pts, refpt, 100.0, 100.0, st.npt,
                                                st.current)
            self.n = n
            pts = self.pts[:n]
            rng, ms = update_compdom(pr, self.grid)
            r = deposit(pts, self.grid, rng, ms)
            ph = solve_openbc(r, ms)
            if rho is not None:
                rho[...] = r
                phi[...] = ph
            w = self.csr_wake(e, z)
            w = w if w is not None else self.wake()
            wk = None
            if w is not None:
                wk = wakefield(self.c, pts, rng, self.grid[2], st.current, st.npt, -refpt[5],
                               st.charge, w)
            cvbkforth1st(self.c, pts, refpt)
            if wk is not None:
                kick1wake(self.c, pts, ph, rng, ms, refpt, st.current, 2 * tau1, st.mass,
                          st.charge, wk)
            else:
                kick1(self.c, pts, ph, rng, ms, refpt, st.current, 2 * tau1, st.mass,
                      st.charge)
        if st.bcurr > 1.0e-15:
            if self.isr_rng is None:
                raise ValueError("a bend with current needs ISR deviates (set isr_rng)")
            rd = self.isr_rng.standard_normal(self.n0)
            isr_chicane(self.c, pts, refpt, 2 * tau1, e.length, e.values, st.mass,
                        rd[pts[:, 8].astype(np.int64) - 1])
        st.z = sector_dipole(self.c, pts, refpt, z, tau1, e.length, e.values, st.charge,
                             st.mass)
        if self._left == 0:
            self._exit = e
        return rho, phi

    def finish_element(self):
        """The exit face of a bend whose last step has run: the run loop
        applies it after that step's diagnostics, at the element's end."""
        e, self._exit = getattr(self, "_exit", None), None
        if e is not None:
            dipole_face(self.c, self.pts[: self.n], np.ctypeslib.as_array(self.st.refpt),
                        e.length, e.values, True, self.st.charge, self.st.mass)

    def particles(self):
        self.finish_element()
        return self.pts[: self.n]

    def diag(self):
        s, m, r2 = diag_sums(self.pts[: self.n])
        refpt = np.array(self.st.refpt[:])
        return diag_row(self.c, s, m, r2, self.st.npt, refpt, self.st.mass)
