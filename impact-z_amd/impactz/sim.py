"""Run a linear-element lattice on the GPU in the reference's segment order.

This is the host loop of run_AccSimulator (src/Contrl/AccSimulator.f90:742-1504)
restricted to Flagmap=1, Flagbc=1, Flagsubstep=0, Flagerr=0 and element types
0 (drift), 1 (quadrupole), 2 (constant focusing) and 3 (solenoid); the per-step procedures all
run on the device through the C ABI.
"""
from __future__ import annotations

import math

import numpy as np

from . import hip
from .beam import Beam, Element

CLIGHT = 299792458.0
SC_CURRENT = float(np.float32(1.0e-10))   # the loop's REAL(4) space-charge threshold


def diag_row(sums, amax, r2max, npt, refpt, mass, scxl, pi):
    """Derived columns of diagnostic1_Output (src/Contrl/Output.f90:247-339).

    Returns 41 values in the order written to fort.18 (5 after z), fort.24,
    fort.25, fort.26, fort.27, fort.29 and fort.30 (6 each).
    """
    t = np.asarray(sums, dtype=np.float64)
    qmc, xl, xt = mass / 1.0e6, scxl, 180.0 / pi
    third = float(np.float32(1.0) / np.float32(3.0))  # REAL(4) exponent (:281)
    den1 = 1.0 / float(npt)
    gam = -refpt[5]
    gambet = math.sqrt(gam * gam - 1.0)
    bet = math.sqrt(gam * gam - 1.0) / gam
    energy = (gam - 1.0) * qmc
    m = t[:6] * den1                      # x0 px0 y0 py0 z0 pz0
    sq = t[6:12] * den1                   # <x^2> ...
    var = sq - m * m
    cross = t[12:15] * den1 - m[0::2] * m[1::2]   # xpx, ypy, zpz
    cub = t[[15, 17, 19, 21, 23, 25]] * den1
    fth = t[[16, 18, 20, 22, 24, 26]] * den1
    m3 = np.abs(cub - 3 * sq * m + 2 * m * m * m) ** third
    m4 = np.sqrt(np.sqrt(np.abs(fth - 4 * cub * m + 6 * sq * m * m - 3 * m ** 4)))
    rms = np.sqrt(var)
    eps = np.sqrt(np.maximum(var[0::2] * var[1::2] - cross * cross, 0.0))
    sc_pos = np.array([xl, xl, xt])       # x, y, z position scales
    sc_mom = np.array([1 / gambet, 1 / gambet, qmc])
    out = [refpt[4], gam, energy, bet, math.sqrt(r2max) * xl]
    for a in range(3):
        emit = eps[a] * (xl if a < 2 else qmc * xt)
        out += [m[2 * a] * sc_pos[a], rms[2 * a] * sc_pos[a],
                m[2 * a + 1] * sc_mom[a], rms[2 * a + 1] * sc_mom[a],
                -cross[a] / eps[a] if eps[a] != 0 else np.nan, emit]
    out += [amax[0] * xl, amax[1] / gambet, amax[2] * xl, amax[3] / gambet,
            amax[4] * xt, amax[5] * qmc]
    for mm in (m3, m4):
        out += [mm[0] * xl, mm[1] / gambet, mm[2] * xl, mm[3] / gambet,
                mm[4] * xt, mm[5] * qmc]
    return np.array(out)


class Simulation:
    """A device-resident bunch + field marched through a lattice."""

    def __init__(self, pts: np.ndarray, beam: Beam, lattice: list[Element],
                 grid: tuple[int, int, int], precision: int = hip.FP64, flagmap: int = 1,
                 qmcclist=None, nptlist=None):
        self.beam = beam
        # charge states of chgupdate_BeamBunch: their q/mc^2 and particle
        # counts, which start as the deck's nptlist (src/Contrl/
        # AccSimulator.f90:155-161) and are recounted after every conv1st
        self.qmcclist = np.asarray(qmcclist if qmcclist is not None
                                   else [beam.charge / beam.mass], dtype=np.float64)
        self.nptlist = np.asarray(nptlist if nptlist is not None else [pts.shape[0]],
                                  dtype=np.int64)
        qm = np.asarray(pts)[:, 6] if len(pts) else np.zeros(0)
        # one q/mc^2 for every particle: the counts follow Npt without a scan
        self._qm_uniform = None if qm.size == 0 or np.any(qm != qm[0]) else float(qm[0])
        self.flagmap = flagmap
        self._left = 0      # segments left in the current element
        self.zedge = 0.0    # z of the current element's entrance
        self._rf = {}       # RF cavity fields on the device library's side
        # read3_Data has filled the module's Fourier coefficient array (Fcoef)
        self._fcoef = False
        self.wk = {"on": 0, "read": 0, "btw": 0, "scale": 1.0, "aa": 0.0, "gg": 0.0,
                   "leng": 0.0, "table": None}
        # the loop's CSR bookkeeping (src/Contrl/AccSimulator.f90:722-740,
        # 1032-1056, 1131-1138, 1374-1385, 1498-1499): zbleng, sentr,
        # flagbendcsr, bendlencsr and the last bend's radius r0
        self.csr = {"zbleng": 0.0, "sentr": 0.0, "flagbend": 0, "bendlen": 0.0, "r0": 0.0}
        # ISRchicane's deviates in bends with current: the reference draws
        # them from the compiler's random_number (normVec1), which no other
        # program reproduces; here a generator the caller seeds gives one
        # standard normal per particle (by upload row) per bend step
        self.isr_rng = None
        self._isr_buf = None
        self.n0 = int(np.asarray(pts).shape[0])
        self.lattice = lattice
        self.grid = tuple(grid)
        self.const = hip.physconst(beam.freq)
        self.bunch = hip.BeamBunch.from_beam(pts, beam, precision)
        self.field = hip.FieldQuant(*self.grid, precision=precision)
        self.dom = hip.CompDom(*self.grid)
        self.z = 0.0
        self.nstep = 0

    def segments(self):
        """The lattice's steps; a -41 element switches the read-in wake as
        the loop reaches it."""
        for e in self.lattice:
            if e.type == -41:
                self.wake_switch(e)
                continue
            for _ in range(e.nseg):
                yield e

    # --- the run loop's wake switches (src/Contrl/AccSimulator.f90:656-720,
    # 759, 859-875, 960-980): flagwake, flagwakeread, flagbtw, scwk, the last
    # SolRF/TWS element's aawk, ggwk, lengwk and the read-in table
    def wake_switch(self, e: Element) -> None:
        """A -41 element: v1 = scale, v2 = file of the wake functions (its
        table is e.field), v3 > 0 on (< 10: no transverse wake), else off."""
        w = self.wk
        if not w["read"]:
            w["on"] = 0
        v = e.values
        if v[1] > 0.0:
            w["table"] = hip.RFData(e.field, fourier=False)
        if v[2] > 0.0:
            w.update(on=1, read=1, scale=v[0])
            if v[2] < 10.0:
                w["btw"] = 2
        else:
            w.update(on=0, read=0, scale=1.0)

    def _wake_entry(self, e: Element) -> None:
        w = self.wk
        if not w["read"]:
            w["on"] = 0
        if e.type in (105, 106):
            v = e.values
            w["aa"], w["gg"], w["leng"] = v[11], v[12], v[13]
            w["on"] = 1 if v[13] > 0.0 else 0

    def _csr_entry(self, e: Element) -> None:
        """Entering a bend (src/Contrl/AccSimulator.f90:1032-1056,1066-1091):
        the transient-drift flag (current and v3 > 500), the bend's length
        for the elements after it, sentr reset, and r0 = |1/hd0|."""
        if e.type != 4:
            return
        v = e.values
        if self.beam.current > 0.0 and v[2] > 500:
            self.csr.update(flagbend=1, bendlen=e.length, sentr=0.0)
        else:
            self.csr["flagbend"] = 0
        hd0 = v[0] / e.length
        self.csr["r0"] = abs(1.0 / hd0)

    def _csr_exit(self, e: Element) -> None:
        """After an element's last step (:1498-1499)."""
        self.csr["zbleng"] += e.length
        self.csr["sentr"] += e.length

    def csr_wake(self, e: Element, z: float):
        """The CSR izh_wake of a step of e whose first half push left z: in a
        bend with v3 > 200 its own steady-state and transient wake, after a
        bend with v3 > 500 the transient wake in every element (flagcsr,
        zwkmin and bendlen of :1131-1138,1374-1385); else None."""
        k = self.csr
        if self.beam.current < SC_CURRENT:
            return None
        if e.type == 4:
            if not e.values[2] > 200:
                return None
            return hip.wake(3, r0=k["r0"], zoff=z - k["zbleng"], blength=e.length)
        if not k["flagbend"]:
            return None
        return hip.wake(3, r0=k["r0"], zoff=z - k["zbleng"] + k["sentr"],
                        blength=k["bendlen"])

    def wake(self):
        """The izh_wake of the current step, or None."""
        w = self.wk
        if not w["on"]:
            return None
        if w["read"]:
            return hip.wake(2, table=w["table"], flagbtw=w["btw"], scale=w["scale"])
        return hip.wake(1, w["aa"], w["gg"], w["leng"], scale=w["scale"])

    def rfdata(self, e: Element):
        """The device library's copy of an RF cavity's field (one per field
        array), or None."""
        if e.field is None:
            return None
        # read3_Data's coefficients for SolRF and for every cavity under the
        # Lorentz integrator (src/Contrl/AccSimulator.f90:826-847)
        fourier = e.type == 105 or self.flagmap != 1
        self._fcoef = self._fcoef or fourier
        key = (id(e.field), fourier)
        if key not in self._rf:
            self._rf[key] = (hip.RFData(e.field, fourier=fourier), e.field)
        return self._rf[key][0]

    def _enter(self, e: Element) -> float:
        """The zedge of the next segment of e: z at the element's entrance
        (src/Contrl/AccSimulator.f90:1025-1026); entering an element also
        sets the wake switches."""
        if self._left == 0:
            self._left, self.zedge = e.nseg, self.z
            self._wake_entry(e)
            self._csr_entry(e)
        self._left -= 1
        return self.zedge

    def finish_element(self) -> None:
        """The exit face of a bend whose last step has run; the run loop
        applies it after that step's diagnostics, at the element's end
        (src/Contrl/AccSimulator.f90:1491-1494)."""
        e, self._exit = getattr(self, "_exit", None), None
        if e is not None:
            self.bunch.bpol_dipole(e.length, e.values)

    def bend_step(self, e: Element, want_moments: bool = True):
        """A bend segment (src/Contrl/AccSimulator.f90:1032-1091, 1114-1138,
        1141-1396, 1417-1448): the entrance face on the element's first step,
        a Sector_Dipole half step, with current the space-charge block (pipe
        radius 100 m, the CSR wake when v3 > 200, kick1wake) and ISRchicane,
        then the second half step.  Zero current runs under either
        integrator, current under Flagmap=1; ISR takes its deviates from
        isr_rng (set a seeded numpy Generator).

        Under Flagmap=2 the run loop also applies map2_BeamBunch's Lorentz
        kick in the bend, whose field (getfld_Dipole, src/Appl/Dipole.f90:
        361-414) takes its pole-face geometry from the Fourier coefficient
        array that read3_Data fills for RF cavities (src/DataStruct/
        Data.f90:54, 348-409).  While no cavity has filled it the array is
        zero and the kick adds nothing, so it is left out; after a cavity
        the bend is refused."""
        cur = self.beam.current
        if cur > 1.0e-15 and self.flagmap != 1:
            raise hip.IZHError("a bend with current runs under Flagmap=1 only")
        if self.flagmap != 1 and self._fcoef:
            raise hip.IZHError("a bend under Flagmap=2 after an RF cavity: its Lorentz "
                               "kick would read the cavity's Fourier coefficients as "
                               "pole-face geometry (getfld_Dipole), which is not restated")
        if cur > 1.0e-15 and self.isr_rng is None:
            raise hip.IZHError("a bend with current needs ISR deviates: set "
                               "Simulation.isr_rng to a seeded numpy Generator")
        entering = self._left == 0
        self._enter(e)
        b = self.bunch
        if entering:
            b.fpol_dipole(e.length, e.values)
        tau1 = 0.5 * e.length / e.nseg
        tau2 = 2.0 * tau1
        self.z = b.sector_dipole(self.z, tau1, e.length, e.values)
        if cur >= SC_CURRENT:
            ptrange, _, _ = b.conv1st(tau2, 100.0, 100.0)
            self.dom.update(ptrange)
            b.charge(self.dom, self.field)
            self.field.update3O(self.dom)
            w = self.csr_wake(e, self.z)
            w = w if w is not None else self.wake()
            if w is not None:
                b.wakefield(self.field, self.dom, w)
            b.cvbkforth1st()
            if w is not None:
                b.kick1wake(tau2, self.field, self.dom)
            else:
                b.kick1(tau2, self.field, self.dom)
        if cur > 1.0e-15:
            if self._isr_buf is None:
                self._isr_buf = hip.DeviceDoubles(self.n0)
            self._isr_buf.upload(self.isr_rng.standard_normal(self.n0))
            b.isr_chicane(tau2, e.length, e.values, self._isr_buf.ptr.value, self.n0)
        self.z = b.sector_dipole(self.z, tau1, e.length, e.values)
        if self._left == 0:
            self._exit = e
            self._csr_exit(e)
        self.nstep += 1
        if not want_moments:
            return None
        s, m, r2 = b.diagnostic1()
        return np.concatenate([s, m, [r2]])

    def step(self, e: Element, want_moments: bool = True):
        self.finish_element()
        if e.type == 4:
            return self.bend_step(e, want_moments)
        zedge = self._enter(e)
        if self.flagmap == 2:
            self.z, mom = self.bunch.step2(
                self.field, self.dom, self.z, 0.5 * e.length / e.nseg,
                hip.element(e.type, e.values, e.length, zedge, rf=self.rfdata(e)), e.radius,
                e.radius, self.beam.current, want_moments, self.wake())
        else:
            self.z, mom = self.bunch.step(self.field, self.dom, self.z,
                                          self.segment_args(e, zedge, self.z),
                                          self.beam.current, want_moments)
        if self._left == 0:
            self._csr_exit(e)
        self.nstep += 1
        return mom

    def step_granular(self, e: Element):
        """The same segment through the individual procedure entry points."""
        zedge = self._enter(e)
        tau1 = 0.5 * e.length / e.nseg
        tau2 = 2.0 * tau1
        b = self.bunch
        rf = self.rfdata(e)
        self.z = b.map1(self.z, tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
        if self.beam.current >= float(np.float32(1.0e-10)):
            ptrange, _, _ = b.conv1st(tau2, e.radius, e.radius)
            self.dom.update(ptrange)
            b.charge(self.dom, self.field)
            self.field.update3O(self.dom)
            b.cvbkforth1st()
            b.kick1(tau2, self.field, self.dom)
        self.z = b.map1(self.z, tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
        if self._left == 0:
            self._csr_exit(e)
        self.nstep += 1

    def segment_args(self, e: Element, zedge: float = 0.0, z: float | None = None) -> hip.Segment:
        """The izh_segment of one space-charge step of element e entered at
        zedge and starting at z, with the wake the switches give it (the
        CSR wake after a bend takes precedence over the 1D wake, as the loop
        computes it last, :1344-1396)."""
        tau1 = 0.5 * e.length / e.nseg
        w = None
        if self.flagmap == 1:
            w = self.csr_wake(e, (self.z if z is None else z) + tau1)
            w = w if w is not None else self.wake()
        return hip.segment(tau1, e.type, e.values, e.mapstp, e.radius, e.radius,
                           e.length, zedge, self.rfdata(e), w)

    def _window(self, start: int, nsteps: int) -> list:
        """The lattice items of steps start..start+nsteps-1 in order: the
        elements of the steps and, as ("switch", element), the -41 elements
        whose next step lies in the window."""
        out, k = [], 0
        for e in self.lattice:
            if e.type == -41:
                if start <= k < start + nsteps:
                    out.append(("switch", e))
                continue
            for _ in range(e.nseg):
                if start <= k < start + nsteps:
                    out.append(e)
                k += 1
        return out

    def _segments_of(self, elements) -> list:
        """izh_segments of consecutive steps, with the entrance bookkeeping
        and wake switches advanced as the steps will advance them (each step
        moves z by 2 tau1)."""
        out, z = [], self.z
        for e in elements:
            if isinstance(e, tuple):
                self.wake_switch(e[1])
                continue
            if self._left == 0:
                self._left, self.zedge = e.nseg, z
                self._wake_entry(e)
                self._csr_entry(e)
            self._left -= 1
            out.append(self.segment_args(e, self.zedge, z))
            if self._left == 0:
                self._csr_exit(e)
            tau1 = 0.5 * e.length / e.nseg
            z = (z + tau1) + tau1
        return out

    def run_fused(self, nsteps: int, start: int = 0, want_moments: bool = True):
        """Steps start..start+nsteps-1 of the lattice through izh_run.

        Returns the per-step records (nsteps x 47: moments, Npt, refpt(5:6))
        or None."""
        if self.flagmap != 1:
            raise hip.IZHError("izh_run fuses the linear-map steps (Flagmap=1); "
                               "Flagmap=2 runs step by step")
        self.finish_element()
        segs = self._segments_of(self._window(start, nsteps))
        self.z, mom = self.bunch.run(self.field, self.dom, self.z, segs,
                                     self.beam.current, want_moments)
        self.nstep += len(segs)
        return mom

    def run_elements(self, elements, want_moments: bool = True):
        """One space-charge step per entry of ``elements`` through izh_run;
        returns the per-step records (47 values: moments, Npt, refpt(5:6))."""
        if self.flagmap != 1:
            raise hip.IZHError("izh_run fuses the linear-map steps (Flagmap=1); "
                               "Flagmap=2 runs step by step")
        self.finish_element()
        elements = list(elements)
        self.z, mom = self.bunch.run(self.field, self.dom, self.z,
                                     self._segments_of(elements), self.beam.current,
                                     want_moments)
        self.nstep += len(elements)
        return mom

    def chg_counts(self, npt: int | None = None) -> np.ndarray:
        """chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728) after a
        conv1st that left npt particles: with one q/mc^2 for all particles
        each state holds all of them or none; otherwise the device counts."""
        if self._qm_uniform is not None or (npt == 0):
            n = self.bunch.info().npt if npt is None else npt
            q = self._qm_uniform if self._qm_uniform is not None else 0.0
            eps = float(np.float32(1.0e-4))
            return np.array([n if abs((q - c) / c) < eps else 0 for c in self.qmcclist],
                            dtype=np.int64)
        return self.bunch.chgupdate(self.qmcclist)

    @property
    def chg_needs_scan(self) -> bool:
        """Charge-state counts need a pass over the particles (q/mc^2 varies)."""
        return self._qm_uniform is None

    def update_chg(self, npt: int | None = None) -> np.ndarray:
        """The counts after a space-charge step (conv1st runs only with
        current, src/Contrl/AccSimulator.f90:1124-1130)."""
        if self.beam.current >= float(np.float32(1.0e-10)):
            self.nptlist = self.chg_counts(npt)
        return self.nptlist

    def diag(self, mom=None, refpt=None):
        """diagnostic1_Output row of the current state, or of a moments record
        (44 values, or 47 with Npt and refpt(5:6) as izh_run gives them)."""
        if mom is None:
            s, m, r2 = self.bunch.diagnostic1()
        else:
            s, m, r2 = mom[:37], mom[37:43], mom[43]
        info = self.bunch.info()
        npt = int(mom[44]) if mom is not None and len(mom) > 44 else info.npt
        rp = np.array(info.refpt[:]) if refpt is None else np.asarray(refpt, dtype=float)
        if refpt is None and mom is not None and len(mom) > 46:
            rp[4], rp[5] = mom[45], mom[46]
        return diag_row(s, m, r2, npt, rp, info.mass, self.const.scxl,
                        self.const.pi)

    def run(self, nsteps: int | None = None, every: int = 1):
        """March through the lattice; returns the diagnostic rows (z first)."""
        rows = [np.concatenate([[self.z], self.diag()])]
        for k, e in enumerate(self.segments()):
            if nsteps is not None and k >= nsteps:
                break
            mom = self.step(e, want_moments=(k + 1) % every == 0)
            if mom is not None:
                rows.append(np.concatenate([[self.z], self.diag(mom)]))
        return np.array(rows)

    def close(self):
        if self._isr_buf is not None:
            self._isr_buf.close()
            self._isr_buf = None
        self.bunch.close()
        self.field.close()
