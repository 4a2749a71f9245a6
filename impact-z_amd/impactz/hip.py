"""ctypes binding of libimpactz_hip.so and a host-side mirror of the
reference's BeamBunch / CompDom / FieldQuant call surface.

Method names follow the reference procedures they replace
(src/Appl/BeamBunch.f90, src/Appl/CompDom.f90, src/Appl/Field.f90,
src/Contrl/Output.f90); errors raise ``IZHError`` where the reference would
``print`` and ``stop``.  Nothing here computes physics: every call goes
through the C ABI, and a missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libimpactz_hip.so")

FP64, MIXED = 0, 1
RHO, PHI, EX, EY, EZ = range(5)

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)

# every symbol of include/impactz_hip.h
SYMBOLS = (
    "izh_last_error", "izh_init", "izh_physconst", "izh_bunch_create",
    "izh_bunch_destroy", "izh_bunch_download", "izh_bunch_upload",
    "izh_bunch_get_info", "izh_bunch_set_info", "izh_quad_maplinear",
    "izh_drift1", "izh_map_linear", "izh_conv1st", "izh_update_compdom",
    "izh_field_create", "izh_field_destroy", "izh_charge", "izh_update3O",
    "izh_cvbkforth1st", "izh_kick1", "izh_diagnostic1", "izh_field_download",
    "izh_field_upload_rho", "izh_step", "izh_run", "izh_timer_enable",
    "izh_timer_read", "izh_run_stats", "izh_probe_copy", "izh_element_maplinear",
    "izh_map1", "izh_fpol_dipole", "izh_sector_dipole", "izh_bpol_dipole",
    "izh_isr_chicane", "izh_drift2", "izh_kick2", "izh_step2", "izh_rfdata_table",
    "izh_rfdata_fourier", "izh_rfdata_destroy", "izh_wakefield", "izh_kick1wake",
    "izh_wake_download", "izh_kick2wake", "izh_chgupdate", "izh_bunch_set_tuning",
    "izh_bunch_get_tuning", "izh_job_create", "izh_job_advance",
    "izh_job_query", "izh_job_finish", "izh_job_info", "izh_job_destroy",
    "izh_bunch_create_from", "izh_diagnostic1_rows",
)

# timer slots (IZH_T_*), named after the reference's Timer accumulators
TIMERS = ("push", "conv1st", "deposit", "green", "fft", "efield", "kick",
          "cvbk", "diag", "pass", "final", "wake")


class IZHError(RuntimeError):
    pass


class Const(C.Structure):
    _fields_ = [("clight", C.c_double), ("pi", C.c_double),
                ("scxl", C.c_double), ("scfreq", C.c_double)]


class BunchInfo(C.Structure):
    _fields_ = [("current", C.c_double), ("mass", C.c_double),
                ("charge", C.c_double), ("npt", C.c_int64),
                ("nptlocal", C.c_int64), ("refpt", C.c_double * 6),
                ("freq", C.c_double), ("precision", C.c_int)]


NVAL = 14          # IZH_NVAL: deck values v1..v14 of an element
RUN_RECORD = 47    # IZH_RUN_RECORD: 44 diagnostic values, Npt, refpt(5:6)


class ElementC(C.Structure):
    """izh_element: type, RK6 steps, values v1..v14 (the deck line's, i.e.
    Param(2:15)), length, entrance z and the RF cavity's field."""
    _fields_ = [("bitype", C.c_int), ("mapstp", C.c_int), ("v", C.c_double * NVAL),
                ("length", C.c_double), ("zedge", C.c_double), ("rf", C.c_void_p)]


class WakeC(C.Structure):
    """izh_wake: kind 1 analytic (iris radius aa, gap gg, period leng), kind 2
    a read-in table (an RFData table of z, Wz, Wx, Wy) with flagbtw; scale =
    scwk; kind 3 the CSR wake of a bend (radius r0, length blength, offset
    zoff of the first mesh line from the bend entrance)."""
    _fields_ = [("kind", C.c_int), ("aa", C.c_double), ("gg", C.c_double),
                ("leng", C.c_double), ("table", C.c_void_p), ("flagbtw", C.c_int),
                ("scale", C.c_double), ("r0", C.c_double), ("zoff", C.c_double),
                ("blength", C.c_double)]


def wake(kind: int, aa: float = 0.0, gg: float = 0.0, leng: float = 0.0,
         table: "RFData | None" = None, flagbtw: int = 0, scale: float = 1.0,
         r0: float = 0.0, zoff: float = 0.0, blength: float = 0.0) -> WakeC:
    w = WakeC()
    w.kind, w.aa, w.gg, w.leng, w.flagbtw, w.scale = kind, aa, gg, leng, flagbtw, scale
    w.r0, w.zoff, w.blength = r0, zoff, blength
    w.table = table.h if table is not None else None
    w._keep = table
    return w


class Segment(C.Structure):
    """izh_segment: an element's half step tau1, the element, the pipe
    aperture and the step's wakefield (None: none)."""
    _fields_ = [("tau1", C.c_double), ("e", ElementC), ("xrad", C.c_double),
                ("yrad", C.c_double), ("wake", C.POINTER(WakeC))]


def values14(v) -> np.ndarray:
    """The element values the library takes (a lone number: a gradient)."""
    out = np.zeros(NVAL)
    if np.isscalar(v):
        out[0] = float(v)
    else:
        vals = np.asarray(v, dtype=np.float64).ravel()[:NVAL]
        out[: vals.size] = vals
    return out


def element(bitype: int, v, length: float = 0.0, zedge: float = 0.0, mapstp: int = 1,
            rf: "RFData | None" = None) -> ElementC:
    e = ElementC()
    e.bitype, e.mapstp = bitype, mapstp
    for k, x in enumerate(values14(v)):
        e.v[k] = x
    e.length, e.zedge = length, zedge
    e.rf = rf.h if rf is not None else None
    e._rf_keep = rf   # the handle lives as long as the element
    return e


def segment(tau1: float, bitype: int, v, mapstp: int, xrad: float, yrad: float,
            length: float = 0.0, zedge: float = 0.0, rf: "RFData | None" = None,
            wake: WakeC | None = None) -> Segment:
    sg = Segment()
    sg.tau1 = tau1
    sg.e = element(bitype, v, length, zedge, mapstp, rf)
    sg._rf_keep = (rf, wake)
    sg.xrad, sg.yrad = xrad, yrad
    if wake is not None:
        sg.wake = C.pointer(wake)
    return sg


class TuningC(C.Structure):
    """izh_tuning: how izh_run works on a bunch (results do not depend on it)."""
    _fields_ = [("resort_period", C.c_int), ("reserved", C.c_int * 7)]


class CompDomC(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("range", C.c_double * 6), ("msize", C.c_double * 3)]


_lib = None


def lib() -> C.CDLL:
    """Load the library (raises if it was not built: no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IZHError(f"{LIB_PATH} not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.izh_last_error.restype = C.c_char_p
    L.izh_init.argtypes = [C.c_int]
    L.izh_physconst.argtypes = [C.c_double, C.POINTER(Const)]
    L.izh_bunch_create.argtypes = [_dp, C.c_int64, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.POINTER(vp)]
    L.izh_bunch_destroy.argtypes = [vp]
    L.izh_bunch_download.argtypes = [vp, _dp, C.c_int64, _i64p]
    L.izh_bunch_upload.argtypes = [vp, _dp, C.c_int64]
    L.izh_bunch_get_info.argtypes = [vp, C.POINTER(BunchInfo)]
    L.izh_bunch_set_info.argtypes = [vp, C.POINTER(BunchInfo)]
    L.izh_quad_maplinear.argtypes = [C.POINTER(Const), C.c_double, C.c_double,
                                     C.c_int, _dp, C.c_double, C.c_double, _dp]
    L.izh_drift1.argtypes = [vp, _dp, C.c_double, C.c_int, C.c_double, C.c_int]
    L.izh_element_maplinear.argtypes = [C.POINTER(Const), C.c_double, C.c_double,
                                        C.POINTER(ElementC), _dp, C.c_double, C.c_double,
                                        _dp]
    L.izh_map1.argtypes = [vp, _dp, C.c_double, C.POINTER(ElementC)]
    L.izh_fpol_dipole.argtypes = [vp, C.POINTER(ElementC)]
    L.izh_bpol_dipole.argtypes = [vp, C.POINTER(ElementC)]
    L.izh_sector_dipole.argtypes = [vp, _dp, C.c_double, C.POINTER(ElementC)]
    L.izh_isr_chicane.argtypes = [vp, C.c_double, C.POINTER(ElementC), vp, C.c_int64]
    L.izh_rfdata_table.argtypes = [_dp, C.c_int, C.POINTER(vp)]
    L.izh_rfdata_fourier.argtypes = [_dp, C.c_int, C.POINTER(vp)]
    L.izh_rfdata_destroy.argtypes = [vp]
    L.izh_wakefield.argtypes = [vp, vp, C.POINTER(CompDomC), C.POINTER(WakeC)]
    L.izh_kick1wake.argtypes = [vp, C.c_double, vp, C.POINTER(CompDomC), C.c_int]
    L.izh_wake_download.argtypes = [vp, _dp, _dp, _dp]
    L.izh_drift2.argtypes = [vp, _dp, C.c_double]
    L.izh_kick2.argtypes = [vp, C.c_double, C.c_double, C.POINTER(ElementC), vp,
                            C.POINTER(CompDomC), C.c_int]
    L.izh_step2.argtypes = [vp, vp, C.POINTER(CompDomC), _dp, C.c_double,
                            C.POINTER(ElementC), C.c_double, C.c_double, C.c_double,
                            C.POINTER(WakeC), _dp]
    L.izh_kick2wake.argtypes = [vp, C.c_double, C.c_double, C.POINTER(ElementC), vp,
                                C.POINTER(CompDomC), C.c_int]
    L.izh_map_linear.argtypes = [vp, _dp]
    L.izh_conv1st.argtypes = [vp, C.c_double, _dp, _i64p, _i64p, C.c_double,
                              C.c_double, C.c_int]
    L.izh_update_compdom.argtypes = [C.POINTER(CompDomC), _dp, C.c_int]
    L.izh_field_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(vp)]
    L.izh_field_destroy.argtypes = [vp]
    L.izh_charge.argtypes = [vp, C.POINTER(CompDomC), vp]
    L.izh_update3O.argtypes = [vp, C.POINTER(CompDomC)]
    L.izh_cvbkforth1st.argtypes = [vp]
    L.izh_kick1.argtypes = [vp, C.c_double, vp, C.POINTER(CompDomC), C.c_int]
    L.izh_diagnostic1.argtypes = [vp, _dp, _dp, _dp]
    L.izh_diagnostic1_rows.argtypes = [vp, C.c_int64, C.c_int64, vp, vp]
    L.izh_chgupdate.argtypes = [vp, C.c_int, _dp, _i64p]
    L.izh_job_create.argtypes = [vp, vp, C.POINTER(CompDomC), C.c_double,
                                 C.POINTER(Segment), C.c_int64, C.c_int64, C.c_double,
                                 C.c_int, C.POINTER(vp)]
    L.izh_job_advance.argtypes = [vp, C.c_int64, _i64p]
    L.izh_job_query.argtypes = [vp, C.POINTER(C.c_int)]
    L.izh_job_finish.argtypes = [vp, _dp, _dp]
    L.izh_job_info.argtypes = [vp, _i64p]
    L.izh_job_destroy.argtypes = [vp]
    L.izh_bunch_create_from.argtypes = [vp, C.c_int64, C.c_int64, C.c_double, C.c_int,
                                        C.POINTER(vp)]
    L.izh_bunch_set_tuning.argtypes = [vp, C.POINTER(TuningC)]
    L.izh_bunch_get_tuning.argtypes = [vp, C.POINTER(TuningC)]
    L.izh_field_download.argtypes = [vp, C.c_int, _dp]
    L.izh_field_upload_rho.argtypes = [vp, _dp]
    L.izh_step.argtypes = [vp, vp, C.POINTER(CompDomC), _dp, C.POINTER(Segment),
                           C.c_double, _dp]
    L.izh_run.argtypes = [vp, vp, C.POINTER(CompDomC), _dp, C.POINTER(Segment),
                          C.c_int64, C.c_double, _dp]
    L.izh_run_stats.argtypes = [vp, _i64p]
    L.izh_probe_copy.argtypes = [C.c_int64, _dp]
    L.izh_timer_enable.argtypes = [C.c_int]
    L.izh_timer_read.argtypes = [_dp, _i64p, C.c_int]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise IZHError(f"izh error {rc}: {lib().izh_last_error().decode()}")


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def init(device: int = 0) -> None:
    check(lib().izh_init(device))


def timer_enable(on: bool = True) -> None:
    check(lib().izh_timer_enable(1 if on else 0))


def timer_read(reset: bool = False) -> dict:
    """{slot: (total ms, launches)} of the kernel timers (HIP events)."""
    ms = np.zeros(len(TIMERS))
    n = np.zeros(len(TIMERS), dtype=np.int64)
    check(lib().izh_timer_read(_p(ms), n.ctypes.data_as(_i64p), 1 if reset else 0))
    return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(TIMERS)}


def probe_copy(nbytes: int) -> float:
    """Time (ms) of one plain device copy kernel of nbytes (counter calibration)."""
    ms = C.c_double()
    check(lib().izh_probe_copy(nbytes, C.byref(ms)))
    return ms.value


def physconst(freq: float) -> Const:
    c = Const()
    check(lib().izh_physconst(freq, C.byref(c)))
    return c


def quad_maplinear(c: Const, tau, gradient, mapstp, refpt, charge, mass):
    """maplinear_Quadrupole (src/Appl/Quadrupole.f90:120-200); host only."""
    r = np.ascontiguousarray(refpt, dtype=np.float64).copy()
    xm = np.zeros(36)
    check(lib().izh_quad_maplinear(C.byref(c), tau, gradient, mapstp, _p(r),
                                   charge, mass, _p(xm)))
    return xm.reshape(6, 6), r


def element_maplinear(c: Const, tau, bitype, v, mapstp, refpt, charge, mass, z=0.0,
                      length=0.0, zedge=0.0, rf: "RFData | None" = None):
    """maplinear_BeamLineElem (src/Appl/BeamLineElem.f90:776-825) for bitypes
    1-3 and the RF cavities 101-105 from z; host only.  Returns (xm 6x6,
    refpt after)."""
    r = np.ascontiguousarray(refpt, dtype=np.float64).copy()
    xm = np.zeros(36)
    e = element(bitype, v, length, zedge, mapstp, rf)
    check(lib().izh_element_maplinear(C.byref(c), z, tau, C.byref(e), _p(r), charge, mass,
                                      _p(xm)))
    return xm.reshape(6, 6), r


class DeviceDoubles:
    """A device array of doubles the host fills, for entry points that read
    device memory the caller provides (the deviates of izh_isr_chicane)."""

    _rt = None

    def __init__(self, n: int):
        if DeviceDoubles._rt is None:
            rt = C.CDLL("libamdhip64.so")
            rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            rt.hipFree.argtypes = [C.c_void_p]
            rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DeviceDoubles._rt = rt
        self.n = int(n)
        self.ptr = C.c_void_p()
        if DeviceDoubles._rt.hipMalloc(C.byref(self.ptr), max(self.n, 1) * 8) != 0:
            raise IZHError(f"hipMalloc of {self.n} doubles failed")

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.n,):
            raise IZHError(f"DeviceDoubles.upload: {a.shape} for {self.n} values")
        if DeviceDoubles._rt.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:
            raise IZHError("hipMemcpy to the device failed")

    def close(self) -> None:
        if self.ptr:
            DeviceDoubles._rt.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RFData:
    """izh_rfdata: an RF cavity's on-axis field as the Data module holds it.
    ``field`` is an rfdata file's contents: an (n, 4) table of z, Ez, Ez',
    Ez'' (read1_Data, src/DataStruct/Data.f90:139-215) or, with
    fourier=True, the coefficient vector (read3_Data, :348-409)."""

    def __init__(self, field, fourier: bool):
        a = np.ascontiguousarray(field, dtype=np.float64)
        h = C.c_void_p()
        if fourier:
            a = a.ravel().copy()
            check(lib().izh_rfdata_fourier(_p(a), a.size, C.byref(h)))
        else:
            a = a.reshape(-1, 4).copy()
            check(lib().izh_rfdata_table(_p(a), a.shape[0], C.byref(h)))
        self.h = h.value
        self.fourier = fourier

    @classmethod
    def for_element(cls, e) -> "RFData | None":
        """The field of a beam.Element (None when it has none)."""
        if e.field is None:
            return None
        return cls(e.field, fourier=e.type == 105)

    def close(self):
        if self.h:
            lib().izh_rfdata_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CompDom:
    """CompDom for one process (src/Appl/CompDom.f90:19-32)."""

    def __init__(self, nx: int, ny: int, nz: int):
        self.c = CompDomC()
        self.c.nx, self.c.ny, self.c.nz = nx, ny, nz

    def update(self, ptrange, flagbc: int = 1) -> None:
        """update_CompDom (src/Appl/CompDom.f90:509-823)."""
        pr = np.ascontiguousarray(ptrange, dtype=np.float64)
        check(lib().izh_update_compdom(C.byref(self.c), _p(pr), flagbc))

    @property
    def range(self):
        return np.array(self.c.range[:])

    @property
    def msize(self):
        return np.array(self.c.msize[:])

    def set(self, rng, msize) -> None:
        for k in range(6):
            self.c.range[k] = float(rng[k])
        for k in range(3):
            self.c.msize[k] = float(msize[k])


class FieldQuant:
    """Device-resident FieldQuant + solver workspace (src/Appl/Field.f90:19-100)."""

    def __init__(self, nx: int, ny: int, nz: int, precision: int = FP64):
        self.shape = (nz, ny, nx)
        self.h = C.c_void_p()
        check(lib().izh_field_create(nx, ny, nz, precision, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().izh_field_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update3O(self, dom: CompDom) -> None:
        """update3O_FieldQuant (src/Appl/Field.f90:108-234)."""
        check(lib().izh_update3O(self.h, C.byref(dom.c)))

    def download(self, which: int) -> np.ndarray:
        out = np.zeros(self.shape)
        check(lib().izh_field_download(self.h, which, _p(out)))
        return out

    def wake_download(self) -> np.ndarray:
        """exwake, eywake, ezwake of the last izh_wakefield, (3, nz)."""
        nz = self.shape[0]
        out = np.zeros((3, nz))
        a, b_, c_ = (np.zeros(nz) for _ in range(3))
        check(lib().izh_wake_download(self.h, _p(a), _p(b_), _p(c_)))
        out[0], out[1], out[2] = a, b_, c_
        return out

    def upload_rho(self, rho: np.ndarray) -> None:
        assert rho.shape == self.shape
        check(lib().izh_field_upload_rho(self.h, _p(np.ascontiguousarray(rho, dtype=np.float64))))


class BeamBunch:
    """Device-resident BeamBunch (src/Appl/BeamBunch.f90:23-33)."""

    def __init__(self, pts: np.ndarray, current, kinetic, mass, charge, freq,
                 phase0=0.0, precision: int = FP64):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        assert pts.ndim == 2 and pts.shape[1] == 9
        self.h = C.c_void_p()
        check(lib().izh_bunch_create(_p(pts), pts.shape[0], current, kinetic,
                                     mass, charge, freq, phase0, precision,
                                     C.byref(self.h)))

    @classmethod
    def from_beam(cls, pts, beam, precision: int = FP64):
        return cls(pts, beam.current, beam.kinetic, beam.mass, beam.charge,
                   beam.freq, beam.phase0, precision)

    @classmethod
    def from_pool(cls, pool: "BeamBunch", offset: int, n: int, current: float,
                  precision: int = FP64) -> "BeamBunch":
        """izh_bunch_create_from: the n particles of a resident pool bunch
        from slot offset on, copied on the device, charge weights for n
        particles at the given current."""
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        check(lib().izh_bunch_create_from(pool.h, offset, n, current, precision,
                                          C.byref(self.h)))
        return self

    def close(self):
        if self.h:
            lib().izh_bunch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> BunchInfo:
        i = BunchInfo()
        check(lib().izh_bunch_get_info(self.h, C.byref(i)))
        return i

    @property
    def nptlocal(self) -> int:
        return int(self.info().nptlocal)

    @property
    def refpt(self) -> np.ndarray:
        return np.array(self.info().refpt[:])

    @property
    def current(self) -> float:
        return float(self.info().current)

    def run_stats(self) -> dict:
        """izh_run_stats: deposit chunk counts of the last izh_run."""
        st = np.zeros(4, dtype=np.int64)
        check(lib().izh_run_stats(self.h, st.ctypes.data_as(_i64p)))
        return {"chunks": int(st[0]), "window_spill": int(st[1]),
                "alive": int(st[2]), "tiles": int(st[3])}

    def getpts(self) -> np.ndarray:
        """getpts_BeamBunch: the local particles as (N, 9) (Pts1 transposed view)."""
        n = self.nptlocal
        out = np.zeros((max(n, 1), 9))
        nn = C.c_int64()
        check(lib().izh_bunch_download(self.h, _p(out), out.shape[0], C.byref(nn)))
        return out[: nn.value]

    def set_refpt(self, refpt) -> None:
        """refptcl(6) of the bunch (izh_bunch_set_info)."""
        i = self.info()
        for k in range(6):
            i.refpt[k] = float(refpt[k])
        check(lib().izh_bunch_set_info(self.h, C.byref(i)))

    def setpts(self, pts: np.ndarray) -> None:
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        check(lib().izh_bunch_upload(self.h, _p(pts), pts.shape[0]))

    # --- the reference's per-step procedures -------------------------------
    def drift1(self, z: float, tau: float, bitype: int, gradient: float = 0.0,
               mapstp: int = 1) -> float:
        """map1_BeamBunch = drift1_BeamBunch (src/Appl/BeamBunch.f90:250-465) for a
        drift or constant-gradient quadrupole; returns z."""
        zz = C.c_double(z)
        check(lib().izh_drift1(self.h, C.byref(zz), tau, bitype, gradient, mapstp))
        return zz.value

    def map1(self, z: float, tau: float, bitype: int, v, mapstp: int = 1,
             length: float = 0.0, zedge: float = 0.0, rf: "RFData | None" = None) -> float:
        """drift1_BeamBunch for every element izh_map1 takes; v = the element's
        values v1..v14, zedge its entrance, rf an RF cavity's field.  Returns z."""
        zz = C.c_double(z)
        e = element(bitype, v, length, zedge, mapstp, rf)
        check(lib().izh_map1(self.h, C.byref(zz), tau, C.byref(e)))
        return zz.value

    def fpol_dipole(self, length: float, v) -> None:
        """A bend's entrance: refptcl(5) += L/(Scxl beta0) and Fpol_Dipole."""
        check(lib().izh_fpol_dipole(self.h, C.byref(element(4, v, length))))

    def bpol_dipole(self, length: float, v) -> None:
        """A bend's exit face Bpol_Dipole."""
        check(lib().izh_bpol_dipole(self.h, C.byref(element(4, v, length))))

    def sector_dipole(self, z: float, tau: float, length: float, v) -> float:
        """Sector_Dipole over tau; returns z + tau."""
        zz = C.c_double(z)
        check(lib().izh_sector_dipole(self.h, C.byref(zz), tau, C.byref(element(4, v, length))))
        return zz.value

    def isr_chicane(self, tau: float, length: float, v, rd_ptr: int, nrd: int) -> None:
        """ISRchicane over tau with the nrd device deviates at rd_ptr (one
        double per particle of the upload, by its row there)."""
        check(lib().izh_isr_chicane(self.h, tau, C.byref(element(4, v, length)),
                                    C.c_void_p(rd_ptr), nrd))

    def drift2(self, z: float, tau: float) -> float:
        """drift2_BeamBunch (src/Appl/BeamBunch.f90:467-509): a drift of tau/2; returns z."""
        zz = C.c_double(z)
        check(lib().izh_drift2(self.h, C.byref(zz), tau))
        return zz.value

    def kick2(self, z: float, tau: float, elem: ElementC, field: FieldQuant | None,
              dom: CompDom | None, flagbc: int = 1) -> None:
        """kick2_BeamBunch (src/Appl/BeamBunch.f90:2187-2508)."""
        check(lib().izh_kick2(self.h, z, tau, C.byref(elem), field.h if field else None,
                              C.byref(dom.c) if dom else None, flagbc))

    def step2(self, field: FieldQuant, dom: CompDom, z: float, tau1: float,
              elem: ElementC, xrad: float, yrad: float, bcurr: float,
              want_moments: bool = False, wake: WakeC | None = None):
        """One Flagmap=2 segment of run_AccSimulator (with kick2wake when a
        wake is given)."""
        zz = C.c_double(z)
        mom = np.zeros(44) if want_moments else None
        check(lib().izh_step2(self.h, field.h, C.byref(dom.c), C.byref(zz), tau1,
                              C.byref(elem), xrad, yrad, bcurr,
                              C.byref(wake) if wake is not None else None,
                              _p(mom) if want_moments else None))
        return zz.value, mom

    def kick2wake(self, z: float, tau: float, elem: ElementC, field: FieldQuant,
                  dom: CompDom, flagbc: int = 1) -> None:
        """kick2wake_BeamBunch (src/Appl/BeamBunch.f90:5048-5410)."""
        check(lib().izh_kick2wake(self.h, z, tau, C.byref(elem), field.h, C.byref(dom.c),
                                  flagbc))

    def conv1st(self, tau: float, xrad: float, yrad: float, flagbc: int = 1):
        """conv1st_BeamBunch (src/Appl/BeamBunch.f90:1252-1665); returns (ptrange, nplc, nptot)."""
        pr = np.zeros(6)
        nplc, nptot = C.c_int64(), C.c_int64()
        check(lib().izh_conv1st(self.h, tau, _p(pr), C.byref(nplc),
                                C.byref(nptot), xrad, yrad, flagbc))
        return pr, nplc.value, nptot.value

    def charge(self, dom: CompDom, field: FieldQuant) -> None:
        """charge_BeamBunch (src/Appl/BeamBunch.f90:3535-3630)."""
        check(lib().izh_charge(self.h, C.byref(dom.c), field.h))

    def cvbkforth1st(self) -> None:
        """cvbkforth1st_BeamBunch (src/Appl/BeamBunch.f90:1667-1701)."""
        check(lib().izh_cvbkforth1st(self.h))

    def wakefield(self, field: FieldQuant, dom: CompDom, w: WakeC) -> None:
        """The run loop's wake fields (src/Contrl/AccSimulator.f90:1272-1338)."""
        check(lib().izh_wakefield(self.h, field.h, C.byref(dom.c), C.byref(w)))

    def kick1wake(self, tau: float, field: FieldQuant, dom: CompDom, flagbc: int = 1) -> None:
        """kick1wake_BeamBunch (src/Appl/BeamBunch.f90:4694-5046)."""
        check(lib().izh_kick1wake(self.h, tau, field.h, C.byref(dom.c), flagbc))

    def kick1(self, tau: float, field: FieldQuant, dom: CompDom, flagbc: int = 1) -> None:
        """map2_BeamBunch = kick1_BeamBunch (src/Appl/BeamBunch.f90:1707-2038)."""
        check(lib().izh_kick1(self.h, tau, field.h, C.byref(dom.c), flagbc))

    def diagnostic1(self):
        """diagnostic1_Output reductions (src/Contrl/Output.f90:119-245)."""
        s, m, r2 = np.zeros(37), np.zeros(6), C.c_double()
        check(lib().izh_diagnostic1(self.h, _p(s), _p(m), C.byref(r2)))
        return s, m, r2.value

    @property
    def resort_period(self) -> int:
        """Passes of izh_run between counting-sort scatters (izh_tuning)."""
        t = TuningC()
        check(lib().izh_bunch_get_tuning(self.h, C.byref(t)))
        return t.resort_period

    @resort_period.setter
    def resort_period(self, period: int) -> None:
        t = TuningC()
        t.resort_period = int(period)
        check(lib().izh_bunch_set_tuning(self.h, C.byref(t)))

    def chgupdate(self, qmcclist) -> np.ndarray:
        """chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728): live particles
        of each charge state in qmcclist (q/mc^2 within relative 1e-4)."""
        q = np.ascontiguousarray(qmcclist, dtype=np.float64)
        out = (C.c_int64 * len(q))()
        check(lib().izh_chgupdate(self.h, len(q), _p(q), out))
        return np.array(out[:], dtype=np.int64)

    def step(self, field: FieldQuant, dom: CompDom, z: float, seg: Segment,
             bcurr: float, want_moments: bool = False):
        """One segment of run_AccSimulator (src/Contrl/AccSimulator.f90:1099-1488)."""
        zz = C.c_double(z)
        mom = np.zeros(44) if want_moments else None
        check(lib().izh_step(self.h, field.h, C.byref(dom.c), C.byref(zz), C.byref(seg),
                             bcurr, _p(mom) if want_moments else None))
        return zz.value, mom

    def run(self, field: FieldQuant, dom: CompDom, z: float, segments,
            bcurr: float, want_moments: bool = True):
        """izh_run: consecutive segments as one device-resident pipeline.

        segments: iterable of Segment (see segment()).
        Returns (z, records[nseg, 47] or None): per step the 44 diagnostic
        values of izh_step, then Npt, refpt(5) and refpt(6)."""
        segs = list(segments)
        arr = (Segment * max(len(segs), 1))()
        for k, sg in enumerate(segs):
            arr[k] = sg if isinstance(sg, Segment) else segment(*sg)
        zz = C.c_double(z)
        mom = np.zeros((len(segs), RUN_RECORD)) if want_moments else None
        check(lib().izh_run(self.h, field.h, C.byref(dom.c), C.byref(zz), arr,
                            len(segs), bcurr, _p(mom) if want_moments else None))
        return zz.value, mom


class Job:
    """A lattice-scan job (izh_job): a bunch, its field and domain stepped
    nsteps times through a periodic lattice (segments of one period),
    enqueued without waiting, with captured graphs of whole periods."""

    def __init__(self, bunch: "BeamBunch", field: FieldQuant, dom: CompDom, z: float,
                 period, nsteps: int, bcurr: float, use_graph: bool = True):
        segs = list(period)
        self._segs = (Segment * len(segs))(*segs)
        self.bunch, self.field, self.dom = bunch, field, dom
        self.nsteps = nsteps
        self.h = C.c_void_p()
        check(lib().izh_job_create(bunch.h, field.h, C.byref(dom.c), z, self._segs,
                                   len(segs), nsteps, bcurr, 1 if use_graph else 0,
                                   C.byref(self.h)))

    def advance(self, max_steps: int) -> int:
        """Enqueue up to max_steps more passes; returns the passes left."""
        left = C.c_int64()
        check(lib().izh_job_advance(self.h, max_steps, C.byref(left)))
        return left.value

    def state(self) -> int:
        """0 busy on the GPU, 1 idle with passes left to enqueue, 2 done."""
        d = C.c_int()
        check(lib().izh_job_query(self.h, C.byref(d)))
        return d.value

    def done(self) -> bool:
        return self.state() == 2

    def finish(self):
        """Wait; returns (z, records nsteps x RUN_RECORD)."""
        zz = C.c_double()
        rec = np.zeros((self.nsteps, RUN_RECORD))
        check(lib().izh_job_finish(self.h, C.byref(zz), _p(rec)))
        return zz.value, rec

    def info(self) -> dict:
        a = (C.c_int64 * 4)()
        check(lib().izh_job_info(self.h, a))
        return {"passes": a[0], "block": a[1], "graph": bool(a[2]), "launches": a[3]}

    def close(self):
        if self.h:
            lib().izh_job_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
