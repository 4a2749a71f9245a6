"""Drive the compiled reference dump harness and read what it writes.

Used only by make_fixtures.py (run in a container holding the reference) and
by tests that run the harness when oracle/_ref/ref_harness is present.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
EXE = os.path.join(ROOT, "oracle", "_ref", "ImpactZexe")


def harness_available() -> bool:
    return os.path.exists(HARNESS) and os.access(HARNESS, os.X_OK)


class _Reader:
    def __init__(self, path):
        self.buf = open(path, "rb").read()
        self.off = 0

    def i32(self):
        v = np.frombuffer(self.buf, np.int32, 1, self.off)[0]
        self.off += 4
        return int(v)

    def f64(self, n):
        v = np.frombuffer(self.buf, np.float64, n, self.off).copy()
        self.off += 8 * n
        return v

    def pts(self):
        n = self.i32()
        return self.f64(9 * n).reshape(n, 9)


def read_dump(path: str, grid) -> dict:
    """One dumped step, in the harness's write order (oracle/ref_harness.f90)."""
    nx, ny, nz = grid
    r = _Reader(path)
    d = {}
    d["z0"] = r.f64(1)[0]
    d["refpt0"] = r.f64(6)
    d["current0"] = r.f64(1)[0]
    d["p0"] = r.pts()
    d["p_map1a"] = r.pts()
    d["ptrange"] = r.f64(6)
    d["current"] = r.f64(1)[0]
    d["p_conv1st"] = r.pts()
    d["range"] = r.f64(6)
    d["msize"] = r.f64(3)
    d["rho"] = r.f64(nx * ny * nz).reshape(nz, ny, nx)
    d["phi"] = r.f64(nx * ny * nz).reshape(nz, ny, nx)
    d["wake"] = np.stack([r.f64(nz), r.f64(nz), r.f64(nz)])   # exwake, eywake, ezwake
    d["p_cvbk"] = r.pts()
    d["p_kick"] = r.pts()
    d["p_map1b"] = r.pts()
    d["z"] = r.f64(1)[0]
    d["refpt"] = r.f64(6)
    assert r.off == len(r.buf), (r.off, len(r.buf))
    return d


def read_fort(path: str) -> np.ndarray:
    """A fort.NN diagnostic file (g0 format: Fortran reals without exponent letters)."""
    rows = []
    for line in open(path):
        vals = []
        for tok in line.split():
            t = tok.replace("D", "E")
            # Flang's g0 may print e.g. ".123E-4" or "-.5"; float() handles both
            vals.append(float(t))
        rows.append(vals)
    return np.array(rows)


def run_harness(pts: np.ndarray, beam, lattice, grid, nsteps: int,
                dump_steps=(), workdir: str | None = None, flagmap: int = 1) -> dict:
    """Run the reference on ``pts`` through ``lattice``; return dumps and diagnostics."""
    from impactz import beam as B
    nx, ny, nz = grid
    n = pts.shape[0]
    tmp = workdir or tempfile.mkdtemp(prefix="impz_ref_")
    with open(os.path.join(tmp, "particles.bin"), "wb") as f:
        np.array([n], dtype=np.int32).tofile(f)
        np.ascontiguousarray(pts, dtype=np.float64).tofile(f)
    with open(os.path.join(tmp, "harness.in"), "w") as f:
        f.write(f"{n} {nx} {ny} {nz}\n")
        f.write(f"{beam.current!r} {beam.kinetic!r} {beam.mass!r} "
                f"{beam.charge!r} {beam.freq!r} {beam.phase0!r}\n")
        f.write(f"{flagmap}\n")
        f.write(f"{len(lattice)}\n")
        for e in lattice:
            f.write(f"{e.length!r} {e.nseg} {e.mapstp} {e.type} " +
                    " ".join(repr(float(v)) for v in e.values) + "\n")
            if e.field is not None:
                B.write_rfdata(os.path.join(tmp, B.rfdata_name(B.field_file_id(e))), e.field)
        f.write(f"{nsteps} {len(dump_steps)}\n")
        if dump_steps:
            f.write(" ".join(str(s) for s in dump_steps) + "\n")
        f.write("particles.bin\n")
    subprocess.run(f"ulimit -s unlimited; exec {HARNESS} harness.in",
                   shell=True, cwd=tmp, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)
    out = {"dumps": {}, "dir": tmp}
    for s in dump_steps:
        out["dumps"][s] = read_dump(os.path.join(tmp, f"dump_{s}.bin"), grid)
    for k in (18, 24, 25, 26, 27, 29, 30):
        out[f"fort{k}"] = read_fort(os.path.join(tmp, f"fort.{k}"))
    r = _Reader(os.path.join(tmp, "final.bin"))
    out["final"] = r.pts()
    out["final_z"] = r.f64(1)[0]
    out["final_refpt"] = r.f64(6)
    out["final_current"] = r.f64(1)[0]
    return out
