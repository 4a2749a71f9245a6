"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run from the repository root in a container that holds the reference build
(``make -f oracle/Makefile.ref``):

    python tests/golden/make_fixtures.py

Every fixture is data only: the seeded input particles, the lattice and the
quantities the reference harness (oracle/ref_harness.f90) dumped.  The
reference's own tests hold no golden vectors, so these are the vectors.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "impact-z_amd"))
sys.path.insert(0, ROOT)

from impactz import beam as B  # noqa: E402
from tests.golden import refdump as R  # noqa: E402

STAGES = ("p0", "p_map1a", "p_conv1st", "p_cvbk", "p_kick", "p_map1b")
# every file diagnostic1_Output writes (src/Contrl/Output.f90:330-352)
EXE_UNITS = (18, 24, 25, 26, 27, 28, 29, 30, 32)


def lattice_array(lattice):
    """length, nseg, mapstp, type, then v1..v4 (or v1..v14 when an element
    has more values), one row per element."""
    wide = any(e.more for e in lattice)
    return np.array([[e.length, e.nseg, e.mapstp, e.type,
                      *(e.values if wide else e.values[:4])] for e in lattice])


def fixture(name, n, grid, lattice, nsteps, dump_steps, seed=1,
            beam=B.Beam(), stages=STAGES, flagmap=1):
    pts = B.gaussian_bunch(n, seed=seed, beam=beam)
    ref = R.run_harness(pts, beam, lattice, grid, nsteps, dump_steps, flagmap=flagmap)
    out = {
        "flagmap": np.array(flagmap),
        "pts_in": pts,
        "grid": np.array(grid),
        "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                          beam.freq, beam.phase0]),
        "lattice": lattice_array(lattice),
        "nsteps": np.array(nsteps),
        "dump_steps": np.array(dump_steps),
        "final": ref["final"],
        "final_refpt": ref["final_refpt"],
        "final_current": np.array(ref["final_current"]),
        "final_z": np.array(ref["final_z"]),
    }
    for i, e in enumerate(lattice):
        if e.field is not None:
            out[f"field{i}"] = np.asarray(e.field, dtype=np.float64)
    for k in (18, 24, 25, 26, 27, 29, 30):
        out[f"fort{k}"] = ref[f"fort{k}"]
    for s in dump_steps:
        d = ref["dumps"][s]
        for key in stages:
            out[f"s{s}_{key}"] = d[key]
        for key in ("ptrange", "range", "msize", "rho", "phi", "refpt", "refpt0", "wake"):
            out[f"s{s}_{key}"] = d[key]
        out[f"s{s}_current"] = np.array(d["current"])
        out[f"s{s}_current0"] = np.array(d["current0"])
        out[f"s{s}_z0"] = np.array(d["z0"])
        out[f"s{s}_z"] = np.array(d["z"])
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def cf_sol_lattice():
    """Constant focusing (src/Appl/ConstFoc.f90) and solenoid (src/Appl/Sol.f90)."""
    return [B.Element(0.1, 4, 10, 2, 30.0, 20.0, 5.0, 1.0),
            B.Element(0.2, 4, 10, 0, 1.0),
            B.Element(0.1, 4, 10, 3, 1.5, 0.0, 1.0),
            B.Element(0.2, 4, 10, 0, 1.0)]


def nl_cf_sol_lattice():
    """Constant focusing and a solenoid for the Lorentz integrator; the
    solenoid's Param(4) (its pipe radius, 10 mm) also sets the length of its
    field ramps in getfld_Sol (2 x 10 mm)."""
    return [B.Element(0.1, 4, 10, 2, 30.0, 20.0, 5.0, 1.0),
            B.Element(0.2, 4, 10, 0, 1.0),
            B.Element(0.1, 4, 10, 3, 1.5, 0.0, 0.01),
            B.Element(0.2, 4, 10, 0, 1.0)]


def kquad_lattice():
    """Quadrupoles given by K (file ID < 0: transfmapK_Quadrupole)."""
    return [B.Element(0.1, 4, 10, 1, 5.4, -1.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0),
            B.Element(0.1, 4, 10, 1, -5.4, -1.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0)]


def io_lattice():
    """FODO with a phase-space dump (-2, file 60) and a checkpoint (-7, file
    50) between the first drift and the defocusing quadrupole."""
    return [B.Element(0.1, 4, 10, 1, 10.0, 0.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0),
            B.Element(0.0, 0, 60, -2, 1.0),
            B.Element(0.0, 0, 50, -7, 0.0),
            B.Element(0.1, 4, 10, 1, -10.0, 0.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0)]


def _run_exe(d):
    subprocess.run(f"ulimit -s unlimited; exec {R.EXE}", shell=True, cwd=d, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def io_fixture(n=1000, grid=(16, 16, 16), beam=B.Beam()):
    """Particle I/O of the reference executable itself (ImpactZexe built from
    source): a run that writes a -2 phase dump and a -7 checkpoint, the run
    restarted from that checkpoint (Rstartflg=1), and a run whose particles
    come from a binary fort.40 with the coordinate offsets of readPar_Dist
    (Flagdist=199)."""
    from impactz import particle_io as P
    pts = B.gaussian_bunch(n, seed=7, beam=beam)
    lat = io_lattice()
    tmp = tempfile.mkdtemp(prefix="impz_io_")
    out = {"pts_in": pts, "grid": np.array(grid),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0]),
           "lattice": np.array([[e.length, e.nseg, e.mapstp, e.type, e.v1, e.v2, e.v3,
                                 e.v4] for e in lat])}
    a = os.path.join(tmp, "a")
    os.makedirs(a)
    B.write_particle_in(os.path.join(a, "particle.in"), pts)
    B.write_deck(os.path.join(a, "ImpactZ.in"), n, grid, beam, lat)
    _run_exe(a)
    with open(os.path.join(a, "fort.50"), "rb") as f:
        out["ckpt50"] = np.frombuffer(f.read(), dtype=np.uint8)
    out["phase60"] = np.loadtxt(os.path.join(a, "fort.60"), ndmin=2)
    for k in EXE_UNITS:
        out[f"a_fort{k}"] = R.read_fort(os.path.join(a, f"fort.{k}"))
    b = os.path.join(tmp, "b")
    os.makedirs(b)
    shutil.copy(os.path.join(a, "fort.50"), os.path.join(b, "fort.50"))
    B.write_deck(os.path.join(b, "ImpactZ.in"), n, grid, beam, lat, rstartflg=1)
    _run_exe(b)
    for k in EXE_UNITS:
        out[f"b_fort{k}"] = R.read_fort(os.path.join(b, f"fort.{k}"))
    c = os.path.join(tmp, "c")
    os.makedirs(c)
    dp = [0.001, 40.0, 0.0, 1.0, 1.0, 2.0e-3, -1.0e-4,
          0.001, 0.001, 0.0, 1.0, 1.0, -3.0e-3, 2.0e-4,
          0.1, 0.001, 0.0, 1.0, 1.0, 0.05, 1.0e-4]
    P.write_binary_particles(os.path.join(c, "fort.40"), pts)
    B.write_deck(os.path.join(c, "ImpactZ.in"), n, grid, beam, lat, flagdist=199,
                 distparam=dp)
    _run_exe(c)
    out["c_distparam"] = np.array(dp)
    for k in EXE_UNITS:
        out[f"c_fort{k}"] = R.read_fort(os.path.join(c, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "io_ckpt.npz")
    np.savez_compressed(path, **out)
    print(f"io_ckpt: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def rf_table(length, n=201, z0=0.05):
    """A half-sine gap field sampled for read1_Data: rows z, Ez, Ez', Ez''
    (unit amplitude; the element's Param(2) scales it).  z starts at z0 so
    the reader's shift to 0 is exercised."""
    s = np.linspace(0.0, length, n)
    k = np.pi / length
    return np.stack([s + z0, np.sin(k * s), k * np.cos(k * s), -k * k * np.sin(k * s)],
                    axis=1)


def rf_fourier(length):
    """read3_Data coefficients: count, field start, end and period, then
    a0, a1, b1, a2, b2, a3, b3 (an even count, so the reader pads it)."""
    return np.array([7.0, 0.0, length, length, 0.4, -1.0, 0.05, 0.3, -0.02, -0.05, 0.01])


def rf_lattice():
    """Every RF cavity class on the linear-map path: DTL with its two
    quadrupoles, CCDTL, CCL and SC share one field table (file 1), SolRF
    with Bz0 = 0.4 T takes Fourier coefficients (file 2)."""
    L = 0.2345
    tab, four = rf_table(L), rf_fourier(0.2)
    rf = (5.0e6, 324.0e6)
    return [B.Element(0.1, 4, 10, 0, 1.0),
            B.Element(L, 6, 20, 101, rf[0], rf[1], -30.0, 1.0,
                      (1.0, 0.03, 12.0, 0.03, -12.0), field=tab),
            B.Element(0.1, 2, 10, 0, 1.0),
            B.Element(L, 4, 20, 102, rf[0], rf[1], 20.0, 1.0, (1.0,), field=tab),
            B.Element(L, 4, 20, 103, 0.8 * rf[0], rf[1], -60.0, 1.0, (1.0,), field=tab),
            B.Element(L, 4, 20, 104, rf[0], rf[1], 0.0, 1.0, (1.0,), field=tab),
            B.Element(0.1, 2, 10, 0, 1.0),
            B.Element(0.2, 6, 20, 105, 3.0e6, rf[1], 45.0, 2.0,
                      (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.4), field=four),
            B.Element(0.1, 4, 10, 0, 1.0)]


def wake_table(n=1001, length=0.1):
    """Wake functions for read1wk_Data: rows z, W_z (V/C/m), W_x and W_y
    (V/C/m^2)."""
    s = np.linspace(0.0, length, n)
    return np.stack([s + 0.01, 2.0e14 * np.exp(-s / 0.004) * np.cos(s / 0.01),
                     3.0e18 * (1.0 - np.exp(-s / 0.002)),
                     2.0e18 * (1.0 - np.exp(-s / 0.003))], axis=1)


def wake_file_lattice():
    """A FODO period with the read-in wake of a -41 element (file 3): on with
    its transverse part (flag 20, scale 1.5), then switched to the
    longitudinal part only (flag 5, scale 0.8) before the defocusing half."""
    tab = wake_table()
    return [B.Element(0.0, 0, 1, -41, 1.5, 3.0, 20.0, field=tab),
            B.Element(0.1, 4, 10, 1, 10.0, 0.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0),
            B.Element(0.0, 0, 1, -41, 0.8, 3.0, 5.0, field=tab),
            B.Element(0.1, 4, 10, 1, -10.0, 0.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0)]


def wake_solrf_lattice():
    """SolRF with the analytic short-range wake of its Param(13:15) (iris
    radius, gap, period) around drifts that switch it off again."""
    four = rf_fourier(0.2)
    return [B.Element(0.1, 4, 10, 0, 1.0),
            B.Element(0.2, 8, 20, 105, 3.0e6, 324.0e6, 45.0, 2.0,
                      (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.4, 10.75e-3, 30.33e-3, 33.321e-3),
                      field=four),
            B.Element(0.1, 4, 10, 0, 1.0)]


def exe_wake_fixture(n=1000, grid=(16, 16, 32), beam=B.Beam()):
    """The reference executable (ImpactZexe built from source) on a deck with
    the -41 wake elements of wake_file_lattice: fort.18/24-26 of its own run
    loop, wake switches included."""
    pts = B.gaussian_bunch(n, seed=5, beam=beam)
    lat = wake_file_lattice()
    tmp = tempfile.mkdtemp(prefix="impz_wk_")
    B.write_particle_in(os.path.join(tmp, "particle.in"), pts)
    B.write_deck(os.path.join(tmp, "ImpactZ.in"), n, grid, beam, lat)
    _run_exe(tmp)
    out = {"pts_in": pts, "grid": np.array(grid),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0]),
           "lattice": lattice_array(lat)}
    for i, e in enumerate(lat):
        if e.field is not None:
            out[f"field{i}"] = np.asarray(e.field, dtype=np.float64)
    for k in EXE_UNITS:
        out[f"fort{k}"] = R.read_fort(os.path.join(tmp, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "exe_wake.npz")
    np.savez_compressed(path, **out)
    print(f"exe_wake: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def nl_rf_lattice():
    """The DTL (with its quadrupoles) and a CCL under the Lorentz integrator:
    their fields from read3_Data's coefficients (a0, then cosine and sine
    pairs over the element length)."""
    coef = np.array([0.2, 1.0, 0.0, -0.3, 0.05, 0.1, -0.02])
    L = 0.1
    return [B.Element(0.05, 4, 10, 0, 1.0),
            B.Element(L, 12, 20, 101, 2.5e6, 324.0e6, 222.0, 7.0,
                      (0.0065, 0.02, 60.0, 0.02, -60.0), field=coef),
            B.Element(0.05, 4, 10, 0, 1.0),
            B.Element(L, 8, 20, 103, 2.0e6, 324.0e6, 150.0, 7.0, (0.0065,), field=coef)]


def nl_solrf_lattice():
    """SolRF under the Lorentz integrator: its coefficient file holds the RF
    section (count 7, over the whole element) and the solenoid section
    (count 5, over its first 0.15 m), as getfld_SolRF reads them."""
    coef = np.array([7.0, 0.0, 0.2, 0.2, 0.4, -1.0, 0.05, 0.3, -0.02, -0.05, 0.01,
                     5.0, 0.0, 0.15, 0.15, 1.2, 0.6, 0.0, -0.1, 0.02])
    return [B.Element(0.05, 4, 10, 0, 1.0),
            B.Element(0.2, 12, 20, 105, 3.0e6, 324.0e6, 45.0, 4.0,
                      (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.3), field=coef),
            B.Element(0.05, 4, 10, 0, 1.0)]


def nl_wake_solrf_lattice():
    """SolRF with its analytic short-range wake under the Lorentz integrator
    (kick2wake)."""
    lat = nl_solrf_lattice()
    e = lat[1]
    lat[1] = B.Element(e.length, e.nseg, e.mapstp, e.type, e.v1, e.v2, e.v3, e.v4,
                       (1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.3, 10.75e-3, 30.33e-3, 33.321e-3),
                       field=e.field)
    return lat


def example1_fixture():
    """The reference's own examples/Example1 (H- through nine DTL cells,
    two quadrupoles and a drift, Lorentz integrator, 200 particles) run by
    ImpactZexe built from source.  Its inputs are kept as data: the
    particles of its particle.in, the element values and the Fourier
    coefficients of its rfdataN.in files."""
    from impactz.deck import lattice_of, read_deck, read_particle_in, read_rfdata
    ex = "/root/reference/examples/Example1"
    deck = read_deck(os.path.join(ex, "ImpactZ.in"))
    lat, _ = lattice_of(deck)
    fields = {}
    for i, e in enumerate(lat):
        if 101 <= e.type <= 105:
            fields[i] = read_rfdata(os.path.join(ex, B.rfdata_name(B.field_file_id(e))), 1)
    lat = [B.Element(e.length, e.nseg, e.mapstp, e.type, e.v1, e.v2, e.v3, e.v4, e.more,
                     field=fields.get(i)) for i, e in enumerate(lat)]
    pts = read_particle_in(os.path.join(ex, "particle.in"))
    beam = deck.beam()
    grid = deck.grid
    tmp = tempfile.mkdtemp(prefix="impz_ex1_")
    B.write_particle_in(os.path.join(tmp, "particle.in"), pts)
    B.write_deck(os.path.join(tmp, "ImpactZ.in"), pts.shape[0], grid, beam, lat, flagmap=2)
    _run_exe(tmp)
    out = {"pts_in": pts, "grid": np.array(grid), "flagmap": np.array(2),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0]),
           "lattice": lattice_array(lat)}
    for i, f in fields.items():
        out[f"field{i}"] = f
    for k in EXE_UNITS:
        out[f"fort{k}"] = R.read_fort(os.path.join(tmp, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "example1.npz")
    np.savez_compressed(path, **out)
    print(f"example1: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def switch_lattice():
    """Linear maps through a FODO half-period, a type -25 element switching
    to the Lorentz integrator (Flagmap = its mapstp, 2), then constant
    focusing, a solenoid with 10 mm field ramps and drifts."""
    return [B.Element(0.1, 4, 10, 1, 10.0, 0.0, 1.0),
            B.Element(0.4, 4, 10, 0, 1.0),
            B.Element(0.0, 0, 2, -25, 0.0),
            B.Element(0.1, 4, 10, 2, 30.0, 20.0, 5.0, 1.0),
            B.Element(0.2, 4, 10, 0, 1.0),
            B.Element(0.1, 4, 10, 3, 1.5, 0.0, 0.01),
            B.Element(0.2, 4, 10, 0, 1.0)]


def exe_nonlinear_fixture(n=1000, grid=(16, 16, 16), beam=B.Beam()):
    """Runs of the reference executable (ImpactZexe built from source) with
    the Lorentz integrator: a deck that starts with Flagmap=1 and switches
    to Flagmap=2 at a -25 element, and a Flagmap=2 deck of the FODO period."""
    pts = B.gaussian_bunch(n, seed=3, beam=beam)
    tmp = tempfile.mkdtemp(prefix="impz_nl_")
    out = {"pts_in": pts, "grid": np.array(grid),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0])}
    for tag, lat, flagmap in (("sw", switch_lattice(), 1), ("nl", B.fodo(1), 2)):
        d = os.path.join(tmp, tag)
        os.makedirs(d)
        B.write_particle_in(os.path.join(d, "particle.in"), pts)
        B.write_deck(os.path.join(d, "ImpactZ.in"), n, grid, beam, lat, flagmap=flagmap)
        _run_exe(d)
        out[f"{tag}_flagmap"] = np.array(flagmap)
        out[f"{tag}_lattice"] = np.array([[e.length, e.nseg, e.mapstp, e.type, e.v1, e.v2,
                                           e.v3, e.v4] for e in lat])
        for k in EXE_UNITS:
            out[f"{tag}_fort{k}"] = R.read_fort(os.path.join(d, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "exe_nonlinear.npz")
    np.savez_compressed(path, **out)
    print(f"exe_nonlinear: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def bend_lattice():
    """Two bends at zero current with pole-face angles, curvatures, fringe
    integrals and k1 (which enters the pole faces), between drifts and a
    quadrupole, then a -2 phase dump to fort.60."""
    return [B.Element(0.3, 2, 1, 0, 1.0),
            B.Element(0.8, 4, 1, 4, 0.25, 0.5, 0.0, 0.02,
                      more=(0.1, 0.15, 0.5, -0.3, 0.4)),
            B.Element(0.4, 2, 1, 0, 1.0),
            B.Element(0.1, 2, 10, 1, 5.0, 0.0, 1.0),
            B.Element(0.5, 3, 1, 4, -0.15, 0.0, 0.0, 0.015,
                      more=(-0.05, 0.0, 0.0, 0.0, 0.5)),
            B.Element(0.2, 2, 1, 0, 1.0),
            B.Element(0.0, 0, 60, -2, 1.0)]


def exe_bend_fixture(n=1000, grid=(16, 16, 16)):
    """Bends (type 4) through the reference executable (ImpactZexe built from
    source) at zero current, with the linear-map and the Lorentz integrator:
    the moments every step and the final particles of the -2 dump."""
    beam = B.Beam(current=0.0)
    pts = B.gaussian_bunch(n, seed=11, beam=beam)
    lat = bend_lattice()
    tmp = tempfile.mkdtemp(prefix="impz_bend_")
    out = {"pts_in": pts, "grid": np.array(grid),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0]),
           "lattice": np.array([[e.length, e.nseg, e.mapstp, e.type, *e.values]
                                for e in lat])}
    for flagmap in (1, 2):
        d = os.path.join(tmp, f"m{flagmap}")
        os.makedirs(d)
        B.write_particle_in(os.path.join(d, "particle.in"), pts)
        B.write_deck(os.path.join(d, "ImpactZ.in"), n, grid, beam, lat, flagmap=flagmap)
        _run_exe(d)
        out[f"m{flagmap}_phase60"] = np.loadtxt(os.path.join(d, "fort.60"), ndmin=2)
        for k in EXE_UNITS:
            out[f"m{flagmap}_fort{k}"] = R.read_fort(os.path.join(d, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "exe_bend.npz")
    np.savez_compressed(path, **out)
    print(f"exe_bend: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


def exe_chg_fixture(n=2000, grid=(16, 16, 16)):
    """Two charge states through the reference executable (ImpactZexe built
    from source): half the particles with the beam's q/mc^2 and charge weight
    for 0.06 A, half with half its q/mc^2 and weights for 0.04 A, through two
    FODO periods in a 2.2 mm pipe, so conv1st loses particles of both states
    and chgupdate_BeamBunch recounts them.  Keeps every diagnostic file,
    fort.28 (particle counts) and fort.32 (counts per charge state) too."""
    beam = B.Beam()
    pts = B.gaussian_bunch(n, seed=13, beam=beam)
    na = n // 2
    qa = beam.charge / beam.mass
    cs = [(na, 0.06, qa), (n - na, 0.04, 0.5 * qa)]
    pts[:na, 6], pts[na:, 6] = cs[0][2], cs[1][2]
    pts[:na, 7] = cs[0][1] / beam.freq / na
    pts[na:, 7] = cs[1][1] / beam.freq / (n - na)
    lat = B.fodo(2, radius=0.0022)
    tmp = tempfile.mkdtemp(prefix="impz_chg_")
    B.write_particle_in(os.path.join(tmp, "particle.in"), pts)
    B.write_deck(os.path.join(tmp, "ImpactZ.in"), n, grid, beam, lat, charge_states=cs)
    _run_exe(tmp)
    out = {"pts_in": pts, "grid": np.array(grid),
           "beam": np.array([beam.current, beam.kinetic, beam.mass, beam.charge,
                             beam.freq, beam.phase0]),
           "lattice": lattice_array(lat), "charge_states": np.array(cs)}
    for k in EXE_UNITS:
        out[f"fort{k}"] = R.read_fort(os.path.join(tmp, f"fort.{k}"))
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "exe_chg.npz")
    np.savez_compressed(path, **out)
    print(f"exe_chg: {os.path.getsize(path) / 1024:.0f} KiB")
    return out


CSR_CASES = [
    # r0 (m), bend length (m), gamma, first line's distance from the entrance
    # (m), line spacing (m), lines: every regime of csrwakeTrIGF_FieldQuant
    (10.0, 0.5, 100.0, 0.2, 1e-6, 64),       # inside the bend: A and B
    (10.0, 0.5, 100.0, 0.02, 1e-6, 64),      # near the entrance, mostly A
    (10.0, 0.5, 100.0, 0.49997, 1e-6, 64),   # across the exit
    (10.0, 0.5, 100.0, 0.6, 1e-6, 64),       # past the bend: C and D
    (10.0, 0.5, 100.0, 1.5, 1e-6, 64),       # far past: D
    (2.0, 0.3, 40.0, 0.1, 5e-6, 48),
    (2.0, 0.3, 40.0, 0.35, 5e-6, 48),
    (5.0, 0.2, 1.16, 0.05, 2e-4, 32),        # low gamma
    (10.0, 0.5, 100.0, -2e-5, 1e-6, 64),     # first lines before the entrance
]


CSR_META = ("the reference's csrwakeTrIGF_FieldQuant (oracle/_ref/ref_csr) on 9 line "
            "densities: inside the bend (entrance transient, steady state), across the "
            "exit, past it (exit transients), low gamma, lines before the entrance")


def csr_fixture(seed=5):
    """The reference's own csrwakeTrIGF_FieldQuant (run by oracle/_ref/ref_csr)
    on noisy Gaussian line densities for the cases of CSR_CASES."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_csr")
    rng = np.random.default_rng(seed)
    tmp = tempfile.mkdtemp(prefix="impz_csr_")
    rows, rhos = [], []
    with open(os.path.join(tmp, "csr.in"), "w") as f:
        f.write(f"{len(CSR_CASES)}\n")
        for r0, blen, gam, ptmin, hx, nx in CSR_CASES:
            z = np.arange(nx)
            rho = np.exp(-0.5 * ((z - nx / 2) / (nx / 6)) ** 2) * 1e3 * \
                (1 + 0.1 * rng.standard_normal(nx))
            f.write(f"{nx} {r0!r} {ptmin!r} {hx!r} {blen!r} {gam!r}\n")
            f.write(" ".join(repr(float(v)) for v in rho) + "\n")
            rows.append([nx, r0, ptmin, hx, blen, gam])
            rhos.append(rho)
    subprocess.run([exe, "csr.in", "csr.out"], cwd=tmp, check=True, stdin=subprocess.DEVNULL,
                   timeout=120)
    ez = np.fromfile(os.path.join(tmp, "csr.out"), dtype=np.float64)
    shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, "csr.npz")
    np.savez_compressed(path, cases=np.array(rows), rho=np.concatenate(rhos), ez=ez)
    print(f"csr: {os.path.getsize(path) / 1024:.0f} KiB")


def main(only=None):
    """Write every fixture, or only the named ones (by file name without
    .npz; the executable runs io_ckpt, example1, exe_wake, exe_nonlinear,
    exe_bend, exe_chg and csr by their function)."""
    if only:
        table = {"io_ckpt": io_fixture, "example1": example1_fixture,
                 "exe_wake": exe_wake_fixture, "exe_nonlinear": exe_nonlinear_fixture,
                 "exe_bend": exe_bend_fixture, "exe_chg": exe_chg_fixture,
                 "csr": csr_fixture}
        for name in only:
            table[name]()
        return
    if not R.harness_available():
        raise SystemExit("oracle/_ref/ref_harness missing: make -f oracle/Makefile.ref")
    meta = {}
    # one FODO period on a cubic grid; dumps in QF (step 1) and the drift (step 5)
    fixture("fodo_16cube", 1000, (16, 16, 16), B.fodo(1), 16, (1, 5))
    meta["fodo_16cube"] = "1000 particles seed 1, 16^3, 1 FODO period, dumps at steps 1 (QF) and 5 (drift)"
    # anisotropic grid: catches swapped axes in deposit, FFT and gather
    fixture("fodo_aniso", 1000, (8, 16, 32), B.fodo(1), 6, (2,))
    meta["fodo_aniso"] = "1000 particles seed 1, grid 8x16x32, dump at step 2"
    # lossy pipe: conv1st drops particles and rescales the current
    lossy = B.fodo(1, radius=0.0022)
    fixture("fodo_lossy", 1000, (16, 16, 16), lossy, 4, (1, 2))
    meta["fodo_lossy"] = "pipe radius 2.2 mm: particles lost in conv1st"
    # moment trajectory over two periods (fort.24-30 every step), no dumps
    fixture("fodo_traj", 2000, (16, 16, 32), B.fodo(2), 32, (), stages=())
    meta["fodo_traj"] = "2000 particles, 16x16x32, 32 steps, diagnostics only"
    # the other linear elements: constant focusing, solenoid, K-mode quads
    fixture("cf_sol", 1000, (16, 16, 16), cf_sol_lattice(), 16, (2, 10))
    meta["cf_sol"] = ("constant focusing (type 2), drift, solenoid (type 3), drift; "
                      "dumps at steps 2 (constant focusing) and 10 (solenoid)")
    fixture("kquad", 1000, (16, 16, 16), kquad_lattice(), 16, (3,))
    meta["kquad"] = "FODO of K-mode quadrupoles (file ID -1, K = +-5.4 1/m^2), dump at step 3"
    # the nonlinear Lorentz integrator (Flagmap=2)
    fixture("nl_fodo", 1000, (16, 16, 16), B.fodo(1), 16, (1, 5), flagmap=2)
    meta["nl_fodo"] = "Flagmap=2 through one FODO period, dumps at steps 1 (QF) and 5 (drift)"
    fixture("nl_cf_sol", 1000, (16, 16, 16), nl_cf_sol_lattice(), 16, (2, 10, 11),
            flagmap=2)
    meta["nl_cf_sol"] = ("Flagmap=2: constant focusing, drift, solenoid with 20 mm end "
                         "ramps, drift; dumps at steps 2, 10 and 11")
    io_fixture()
    meta["io_ckpt"] = ("ImpactZexe runs: -2 phase dump and -7 checkpoint after 8 steps, "
                       "the run restarted from that checkpoint, and Flagdist=199 binary input")
    fixture("rf", 1000, (16, 16, 16), rf_lattice(), 36, (6, 22, 30))
    meta["rf"] = ("RF cavities on the linear-map path: DTL (with quadrupoles), CCDTL, CCL, "
                  "SC on a field table (rfdata1.in), SolRF on Fourier coefficients "
                  "(rfdata2.in); dumps at steps 6 (DTL), 22 (SC) and 30 (SolRF)")
    fixture("wake_file", 1000, (16, 16, 32), wake_file_lattice(), 16, (2, 12))
    meta["wake_file"] = ("read-in wake functions of a -41 element (rfdata3.in): transverse "
                         "and longitudinal from step 1 (dump 2), longitudinal only after the "
                         "second -41 (dump 12)")
    fixture("wake_solrf", 1000, (16, 16, 32), wake_solrf_lattice(), 16, (6,))
    meta["wake_solrf"] = "SolRF with its analytic short-range wake (Param(13:15)), dump 6"
    fixture("nl_rf", 1000, (16, 16, 16), nl_rf_lattice(), 28, (6, 22), flagmap=2)
    meta["nl_rf"] = ("Flagmap=2 through a DTL (with quadrupoles) and a CCL on Fourier "
                     "coefficients (rfdata7.in); dumps at steps 6 (DTL) and 22 (CCL)")
    fixture("nl_solrf", 1000, (16, 16, 16), nl_solrf_lattice(), 20, (7,), flagmap=2)
    meta["nl_solrf"] = ("Flagmap=2 through SolRF: RF and solenoid sections of its coefficient "
                        "file (rfdata4.in); dump at step 7")
    fixture("nl_wake_file", 1000, (16, 16, 32), wake_file_lattice(), 16, (6,), flagmap=2)
    meta["nl_wake_file"] = ("Flagmap=2 with the read-in wake of -41 elements (kick2wake), "
                            "dump at step 6")
    fixture("nl_wake_solrf", 1000, (16, 16, 32), nl_wake_solrf_lattice(), 20, (7,),
            flagmap=2)
    meta["nl_wake_solrf"] = "Flagmap=2 through SolRF with its analytic wake, dump at step 7"
    example1_fixture()
    meta["example1"] = ("examples/Example1 of the reference run by ImpactZexe: inputs kept as "
                        "data (particles, element values, Fourier coefficients), fort.18/24-26")
    exe_wake_fixture()
    meta["exe_wake"] = ("ImpactZexe run of the wake_file lattice: -41 read-in wake on with "
                        "its transverse part, then longitudinal only; fort.18/24-26")
    exe_nonlinear_fixture()
    exe_bend_fixture()
    meta["exe_bend"] = ("ImpactZexe runs of two bends at zero current (pole faces, "
                        "fringe, k1) under Flagmap=1 and 2: fort.18/24-26 and the -2 dump")
    exe_chg_fixture()
    meta["exe_chg"] = ("ImpactZexe run of two charge states (q/mc^2 and half of it, "
                       "0.06 A and 0.04 A) through two FODO periods in a 2.2 mm pipe: "
                       "every diagnostic file, fort.28 and fort.32 included")
    csr_fixture()
    meta["csr"] = CSR_META
    meta["exe_nonlinear"] = ("ImpactZexe runs with the Lorentz integrator: Flagmap=1 "
                             "switched to 2 by a -25 element before constant focusing and "
                             "a solenoid, and a Flagmap=2 FODO period")
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
