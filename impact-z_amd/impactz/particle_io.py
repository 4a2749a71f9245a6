"""Particle files of the reference, for the device-resident bunch.

Binary files are Fortran unformatted sequential files, the form the
reference opens them in (``form="unformatted"``): each record is its payload
between two 4-byte little-endian byte counts.

- ``readPar_Dist`` (src/Appl/Distribution.f90:1717-1751), Flagdist=199: the
  particles come from the binary file fort.N, N = distparam(2) (one process):
  a record with Nptlocal (int32), a record with Pts1(9,Nptlocal), and the
  offsets distparam(6,7,13,14,20,21) added to coordinates 1-6.
- ``outpoint_Output`` / ``inpoint_Output`` (src/Contrl/Output.f90:1749-1787,
  1371-1437): the -7 checkpoint and the Rstartflg=1 restart.  Records: z;
  inb, jstp, ibal, nstep (int32); Meshsize(3); SpatRange(6); LcTabnm(2,1,1)
  (int32); LcTabrg(4,1,1); Nptlocal (int32); refptcl(6); Pts1(9,Nptlocal).
- ``phase_Output`` (src/Contrl/Output.f90:1789-1855), the -2 element: a text
  dump of every samplePeriod-th particle, the 9 values of Pts1 per line, or
  for samplePeriod < 0 the six columns x*Scxl, px, y*Scxl, py, -phi*Scxl,
  -pt - refptcl(6).

The particles of a device bunch move through ``BeamBunch.getpts`` /
``setpts`` (izh_bunch_download / izh_bunch_upload), in the reference's order.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np


class FortranFileError(ValueError):
    pass


def read_records(path: str) -> list[bytes]:
    """The records of a Fortran unformatted sequential file."""
    with open(path, "rb") as f:
        buf = f.read()
    out, off = [], 0
    while off < len(buf):
        if off + 4 > len(buf):
            raise FortranFileError(f"{path}: truncated record header at byte {off}")
        (n,) = struct.unpack_from("<i", buf, off)
        end = off + 4 + n
        if n < 0 or end + 4 > len(buf):
            raise FortranFileError(f"{path}: bad record length {n} at byte {off}")
        (n2,) = struct.unpack_from("<i", buf, end)
        if n2 != n:
            raise FortranFileError(f"{path}: record markers {n} and {n2} differ at byte {off}")
        out.append(buf[off + 4:end])
        off = end + 4
    return out


def write_records(path: str, records) -> None:
    with open(path, "wb") as f:
        for r in records:
            b = r.tobytes() if isinstance(r, np.ndarray) else bytes(r)
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(b)))


def _f64(rec: bytes, n: int | None = None, what: str = "") -> np.ndarray:
    a = np.frombuffer(rec, dtype="<f8")
    if n is not None and a.size != n:
        raise FortranFileError(f"{what}: expected {n} doubles, got {a.size}")
    return a.copy()


def _i32(rec: bytes, n: int | None = None, what: str = "") -> np.ndarray:
    a = np.frombuffer(rec, dtype="<i4")
    if n is not None and a.size != n:
        raise FortranFileError(f"{what}: expected {n} integers, got {a.size}")
    return a.copy()


def _pts_record(pts: np.ndarray) -> np.ndarray:
    """Pts1(9,N) in Fortran order: the 9 values of a particle are contiguous."""
    pts = np.ascontiguousarray(pts, dtype="<f8")
    if pts.ndim != 2 or pts.shape[1] != 9:
        raise ValueError("particles must be (N, 9)")
    return pts


# ---------------------------------------------------------------- Flagdist 199

# distparam (1-based) entries readPar_Dist adds to coordinates 1..6
PAR_OFFSETS = (6, 7, 13, 14, 20, 21)


def write_binary_particles(path: str, pts: np.ndarray) -> None:
    """The binary particle file readPar_Dist reads: Nptlocal, then Pts1."""
    p = _pts_record(pts)
    write_records(path, [np.array([p.shape[0]], dtype="<i4"), p])


def read_binary_particles(path: str, distparam=None) -> np.ndarray:
    """readPar_Dist (src/Appl/Distribution.f90:1717-1751): the particles of a
    binary fort.N, with distparam's coordinate offsets added."""
    recs = read_records(path)
    if len(recs) < 2:
        raise FortranFileError(f"{path}: expected 2 records, got {len(recs)}")
    n = int(_i32(recs[0], 1, path)[0])
    pts = _f64(recs[1], 9 * n, path).reshape(n, 9)
    if distparam is not None:
        dp = np.asarray(distparam, dtype=np.float64)
        for c, k in enumerate(PAR_OFFSETS):
            pts[:, c] = pts[:, c] + dp[k - 1]
    return pts


# ------------------------------------------------------------ -7 checkpoint

@dataclass
class Checkpoint:
    """What outpoint_Output writes and inpoint_Output reads back."""
    z: float
    inb: int              # element index the run stopped at
    jstp: int             # segment counter at that point
    ibal: int             # space-charge steps so far
    nstep: int            # steps so far
    msize: np.ndarray     # Meshsize(3) of the last domain update
    range: np.ndarray     # SpatRange(6)
    lctabnm: np.ndarray   # LcTabnm(2,1,1): local mesh counts (z, y)
    lctabrg: np.ndarray   # LcTabrg(4,1,1): local z and y ranges
    refpt: np.ndarray     # refptcl(6)
    pts: np.ndarray = field(repr=False)  # Pts1 as (Nptlocal, 9)


def write_checkpoint(path: str, ck: Checkpoint) -> None:
    """outpoint_Output for one process (src/Contrl/Output.f90:1749-1787)."""
    p = _pts_record(ck.pts)
    write_records(path, [
        np.array([ck.z], dtype="<f8"),
        np.array([ck.inb, ck.jstp, ck.ibal, ck.nstep], dtype="<i4"),
        np.asarray(ck.msize, dtype="<f8").reshape(3),
        np.asarray(ck.range, dtype="<f8").reshape(6),
        np.asarray(ck.lctabnm, dtype="<i4").reshape(2),
        np.asarray(ck.lctabrg, dtype="<f8").reshape(4),
        np.array([p.shape[0]], dtype="<i4"),
        np.asarray(ck.refpt, dtype="<f8").reshape(6),
        p,
    ])


def read_checkpoint(path: str) -> Checkpoint:
    """inpoint_Output for one process (src/Contrl/Output.f90:1371-1437)."""
    r = read_records(path)
    if len(r) != 9:
        raise FortranFileError(f"{path}: a checkpoint has 9 records, found {len(r)}")
    counters = _i32(r[1], 4, path)
    n = int(_i32(r[6], 1, path)[0])
    return Checkpoint(z=float(_f64(r[0], 1, path)[0]), inb=int(counters[0]),
                      jstp=int(counters[1]), ibal=int(counters[2]), nstep=int(counters[3]),
                      msize=_f64(r[2], 3, path), range=_f64(r[3], 6, path),
                      lctabnm=_i32(r[4], 2, path), lctabrg=_f64(r[5], 4, path),
                      refpt=_f64(r[7], 6, path), pts=_f64(r[8], 9 * n, path).reshape(n, 9))


# -------------------------------------------------------------- -2 phase dump

def write_phase(path: str, pts: np.ndarray, sample_period: int, scxl: float,
                refpt6: float) -> None:
    """phase_Output for one process (src/Contrl/Output.f90:1789-1855).  Values
    are written with 17 significant digits (the reference's g0 edit
    descriptor prints the shortest form of the same values)."""
    p = np.asarray(pts, dtype=np.float64)
    sp = 1 if sample_period == 0 else sample_period
    with open(path, "w") as f:
        if sp > 0:
            sel = p[::sp]
            for row in sel:
                f.write(" ".join(repr(float(v)) for v in row) + "\n")
        else:
            sel = p[::-sp]
            cols = np.stack([sel[:, 0] * scxl, sel[:, 1], sel[:, 2] * scxl, sel[:, 3],
                             -sel[:, 4] * scxl, -sel[:, 5] - refpt6], axis=1)
            for row in cols:
                f.write(" ".join(repr(float(v)) for v in row) + "\n")


def read_phase(path: str) -> np.ndarray:
    return np.loadtxt(path, ndmin=2)
