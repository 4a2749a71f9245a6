"""Beam, lattice and synthetic-particle definitions for the space-charge step.

Units follow the reference particle array ``Pts1(9, N)``
(src/Appl/BeamBunch.f90:9-12, src/Appl/Distribution.f90:221-246):

  0  x / Scxl            (Scxl = c / (2 pi f), src/DataStruct/PhysConst.f90:30)
  1  gamma * beta_x
  2  y / Scxl
  3  gamma * beta_y
  4  phi = omega * dt      [rad]
  5  -delta gamma
  6  q / (m c^2)           [1/eV]
  7  I / (f N) * sign(q)   [C per macro-particle]
  8  particle id

The sampler is this project's own (numpy Philox, counter based), so the same
seed gives the same particles on every machine; the reference's
``random_number`` stream is compiler specific, so parity inputs go to the
reference through ``particle.in`` (Flagdist=19, src/Appl/Distribution.f90:1666).
"""
from __future__ import annotations

import dataclasses
import os
import math

import numpy as np

CLIGHT = 299792458.0
PROTON_MASS_EV = 938.272088e6


@dataclasses.dataclass(frozen=True)
class Beam:
    """Deck line ``I W m q f phi0`` (src/Contrl/Input.f90:180)."""

    current: float = 0.1            # A
    kinetic: float = 150.0e6        # eV
    mass: float = PROTON_MASS_EV    # eV
    charge: float = 1.0             # units of e
    freq: float = 324.0e6           # Hz
    phase0: float = 0.0

    @property
    def gamma0(self) -> float:
        return self.kinetic / self.mass + 1.0

    @property
    def gambet0(self) -> float:
        return math.sqrt(self.gamma0 ** 2 - 1.0)

    @property
    def beta0(self) -> float:
        return self.gambet0 / self.gamma0

    @property
    def scxl(self) -> float:
        # src/DataStruct/PhysConst.f90:29-30: Pi = 2 asin(1d0), Scxl = c/(2 pi f)
        return CLIGHT / (2 * (2 * math.asin(1.0)) * self.freq)


NVAL = 14   # deck values v1..v14 of an element line


@dataclasses.dataclass(frozen=True)
class Element:
    """One lattice line ``L nseg mapstp type v1 v2 v3 v4 /`` (src/Contrl/Input.f90:307-314;
    the values become the element's Param(2:5), src/Contrl/AccSimulator.f90:262-320).

    type 0: drift, v1 = pipe radius.
    type 1: quadrupole, v1 = gradient [T/m] (or K [1/m^2] when v2 < 0: the
            K-mode map), v2 = file ID (0: constant gradient), v3 = pipe radius.
    type 2: constant focusing, v1..v3 = kx0^2, ky0^2, kz0^2 [1/m^2],
            v4 = pipe radius.
    type 3: solenoid, v1 = Bz0 [T], v2 = file ID, v3 = pipe radius.
    types 101-104 (DTL, CCDTL, CCL, SC): v1 = field scale, v2 = RF frequency
            [Hz], v3 = phase [deg], v4 = file ID, v5 = pipe radius; the DTL's
            v6..v9 are its quadrupoles' lengths and gradients.
    type 105 (SolRF): v1..v5 as above, v6..v10 misalignments, v11 = Bz0 [T],
            v12..v14 wakefield.

    type -41 (wakefield switch): v1 = scale, v2 = file ID of the wake
            functions, v3 > 0 turns them on (< 10: no transverse wake).

    ``more`` holds v5..v14; ``field`` the element's rfdata file contents:
    an (n, 4) table (types 101-104, read1_Data; type -41, read1wk_Data) or
    the coefficient vector (type 105, read3_Data).
    """

    length: float
    nseg: int
    mapstp: int
    type: int
    v1: float
    v2: float = 0.0
    v3: float = 0.0
    v4: float = 0.0
    more: tuple = ()
    field: object = dataclasses.field(default=None, compare=False, repr=False)

    @property
    def radius(self) -> float:
        # getradius_BeamLineElem (src/Appl/BeamLineElem.f90:500-552)
        if 101 <= self.type <= 105:
            return self.values[4]
        return {0: self.v1, 2: self.v4}.get(self.type, self.v3)

    @property
    def values(self) -> list[float]:
        """v1..v14 (Param(2:15)), zero-padded."""
        out = [self.v1, self.v2, self.v3, self.v4, *(float(x) for x in self.more)]
        return (out + [0.0] * NVAL)[:NVAL]


def fodo(periods: int, gradient: float = 10.0, quad_len: float = 0.1,
         drift_len: float = 0.4, nseg: int = 4, mapstp: int = 10,
         radius: float = 1.0) -> list[Element]:
    """QF + D + QD + D per period; ``nseg`` space-charge kicks per element."""
    cell = [
        Element(quad_len, nseg, mapstp, 1, +gradient, 0.0, radius),
        Element(drift_len, nseg, mapstp, 0, radius),
        Element(quad_len, nseg, mapstp, 1, -gradient, 0.0, radius),
        Element(drift_len, nseg, mapstp, 0, radius),
    ]
    return cell * periods


@dataclasses.dataclass(frozen=True)
class Spread:
    """rms sizes of the uncorrelated 6D beam in physical units."""

    sig_x: float = 1.0e-3       # m
    sig_xp: float = 1.0e-3      # rad
    sig_y: float = 1.0e-3
    sig_yp: float = 1.0e-3
    sig_phi_deg: float = 5.0    # degrees of RF phase
    sig_dw_over_w: float = 1.0e-3  # relative spread of the KINETIC energy


def _truncated_normal(gen: np.random.Generator, n: int, trunc: float) -> np.ndarray:
    """n x 6 standard normals, each coordinate independently cut at |u| <= trunc."""
    out = np.empty((n, 6))
    filled = 0
    while filled < n:
        need = n - filled
        cand = gen.standard_normal((int(need * 1.02) + 64, 6))
        cand = cand[np.all(np.abs(cand) <= trunc, axis=1)][:need]
        out[filled:filled + len(cand)] = cand
        filled += len(cand)
    return out


def _waterbag6d(gen: np.random.Generator, n: int, rms: float) -> np.ndarray:
    """Uniform fill of a 6D ball whose per-coordinate rms is ``rms``.

    For a uniform d-ball of radius R the per-coordinate variance is R^2/(d+2),
    so R = rms * sqrt(8) in 6D.
    """
    d = gen.standard_normal((n, 6))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = (rms * math.sqrt(8.0)) * gen.random(n) ** (1.0 / 6.0)
    return d * r[:, None]


def normalized_to_pts(u: np.ndarray, beam: Beam, spread: Spread,
                      n_total: int | None = None) -> np.ndarray:
    """Map normalized coordinates (rms 1 per axis) to a (N, 9) Pts1 array."""
    n = u.shape[0]
    n_total = n if n_total is None else n_total
    scxl = beam.scxl
    gb = beam.gambet0
    pts = np.empty((n, 9))
    pts[:, 0] = u[:, 0] * spread.sig_x / scxl
    pts[:, 1] = u[:, 1] * spread.sig_xp * gb
    pts[:, 2] = u[:, 2] * spread.sig_y / scxl
    pts[:, 3] = u[:, 3] * spread.sig_yp * gb
    pts[:, 4] = u[:, 4] * math.radians(spread.sig_phi_deg)
    # Pts1(6) = -delta gamma; the kinetic spread dW/W maps to dgamma = dW/(m c^2)
    pts[:, 5] = -u[:, 5] * spread.sig_dw_over_w * beam.kinetic / beam.mass
    pts[:, 6] = beam.charge / beam.mass
    pts[:, 7] = beam.current / beam.freq / n_total * math.copysign(1.0, beam.charge)
    pts[:, 8] = np.arange(1, n + 1, dtype=np.float64)
    return pts


def gaussian_bunch(n: int, seed: int = 1, beam: Beam = Beam(),
                   spread: Spread = Spread(), trunc: float = 3.0) -> np.ndarray:
    """Uncorrelated 6D Gaussian, every coordinate truncated at ``trunc`` sigma."""
    gen = np.random.Generator(np.random.Philox(seed))
    return normalized_to_pts(_truncated_normal(gen, n, trunc), beam, spread)


def core_halo_bunch(n: int, seed: int = 1, beam: Beam = Beam(),
                    spread: Spread = Spread(), halo_frac: float = 0.2,
                    halo_rms: float = 2.0, trunc: float = 3.0) -> np.ndarray:
    """Gaussian core plus a 6D waterbag halo with ``halo_rms`` x the core rms."""
    gen = np.random.Generator(np.random.Philox(seed))
    n_halo = int(round(n * halo_frac))
    n_core = n - n_halo
    u = np.concatenate([_truncated_normal(gen, n_core, trunc),
                        _waterbag6d(gen, n_halo, halo_rms)])
    return normalized_to_pts(u, beam, spread)


def write_particle_in(path: str, pts: np.ndarray) -> None:
    """Text ``particle.in`` read by readin_Dist (src/Appl/Distribution.f90:1677-1705)."""
    with open(path, "w") as f:
        f.write(f"{pts.shape[0]}\n")
        np.savetxt(f, pts, fmt="%.17e")


def write_deck(path: str, n: int, grid: tuple[int, int, int], beam: Beam,
               lattice: list[Element], flagdiag: int = 1, rstartflg: int = 0,
               flagdist: int = 19, distparam=None, flagmap: int = 1,
               charge_states=None) -> None:
    """An ``ImpactZ.in`` for Flagbc=1 (src/Contrl/Input.f90:62-314) and the
    integrator ``flagmap`` (1 linear map, 2 nonlinear Lorentz): particles
    from particle.in (Flagdist=19) or a binary fort.N (199, the file number in
    distparam[1]); rstartflg=1 restarts from a -7 checkpoint.
    charge_states: (particles, current, q/mc^2) of each charge state, the
    deck's nchrg, nptlist, currlist and qmcclist (default: one state, the
    beam's)."""
    nx, ny, nz = grid
    dp = list(distparam) if distparam is not None else [
        0.001, 0.001, 0.0, 1.0, 1.0, 0.0, 0.0,
        0.001, 0.001, 0.0, 1.0, 1.0, 0.0, 0.0,
        0.1, 0.001, 0.0, 1.0, 1.0, 0.0, 0.0]
    cs = list(charge_states) if charge_states else [(n, beam.current,
                                                     beam.charge / beam.mass)]
    lines = [
        "1 1",
        f"6 {n} {flagmap} 0 {flagdiag}",
        f"{nx} {ny} {nz} 1 1.0 1.0 1.0",
        f"{flagdist} {rstartflg} 0 {len(cs)}",
        " ".join(str(int(c[0])) for c in cs),
        " ".join(repr(float(c[1])) for c in cs),
        " ".join(repr(float(c[2])) for c in cs),
        " ".join(repr(float(v)) for v in dp[0:7]),
        " ".join(repr(float(v)) for v in dp[7:14]),
        " ".join(repr(float(v)) for v in dp[14:21]),
        f"{beam.current!r} {beam.kinetic!r} {beam.mass!r} {beam.charge!r} "
        f"{beam.freq!r} {beam.phase0!r}",
    ]
    for e in lattice:
        if e.type == 0:
            lines.append(f"{e.length!r} {e.nseg} {e.mapstp} 0 {e.v1!r} /")
        else:
            vals = e.values if e.more else e.values[:8]
            lines.append(f"{e.length!r} {e.nseg} {e.mapstp} {e.type} " +
                         " ".join(repr(float(v)) for v in vals) + " /")
        if e.field is not None:
            write_rfdata(os.path.join(os.path.dirname(path) or ".",
                                      rfdata_name(field_file_id(e))), e.field)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def field_file_id(e: Element) -> int:
    """The rfdata file number of an element's field: Param(5) = v4 of an RF
    cavity, Param(3) = v2 of a -41 wakefield element
    (src/Contrl/AccSimulator.f90:821-823,962-965)."""
    return int((e.v2 if e.type == -41 else e.v4) + 0.1)


def rfdata_name(file_id: int) -> str:
    """rfdataN.in, N = the element's file ID (src/DataStruct/Data.f90:139-182)."""
    if not 1 <= file_id <= 999:
        raise ValueError(f"file ID {file_id} outside 1..999")
    return f"rfdata{file_id}.in"


def write_rfdata(path: str, field) -> None:
    """An rfdata file: the rows z Ez Ez' Ez'' of a table (n x 4) or one
    Fourier-series value per line."""
    a = np.asarray(field, dtype=np.float64)
    with open(path, "w") as f:
        for row in (a if a.ndim == 2 else a[:, None]):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")
