"""The host deck path: read an ``ImpactZ.in`` deck and its ``particle.in``,
march the bunch through the lattice on the GPU, and write the reference's
per-step diagnostic files.

The deck reader follows in1_Input / in2_Input (src/Contrl/Input.f90:37-384):
records are whitespace separated, lines whose first non-blank character is
``!`` are comments, Fortran ``d`` exponents are accepted and a lattice line
ends at ``/``.  The element table follows the lattice branch of
run_AccSimulator (src/Contrl/AccSimulator.f90:237-536, 742-1504) for the
elements this package runs: drifts (type 0), quadrupoles (type 1: constant
gradient with file ID <= 1e-5, src/Appl/Quadrupole.f90:265-270, or given by
K with file ID < -1e-6, src/Appl/BeamBunch.f90:305-307), constant focusing
(type 2) and solenoids (type 3), under the linear map integrator (Flagmap=1)
or the nonlinear Lorentz integrator (Flagmap=2), and the RF cavities DTL,
CCDTL, CCL, SC and SolRF (types 101-105) with their rfdataN.in fields under
Flagmap=1 and Flagmap=2; a type -25 element switches
the integrator to its mapstp for the elements after it
(src/Contrl/AccSimulator.f90:807-809).  Zero-length diagnostic elements that
do not change the particles (types -3 to -6, -8) are skipped; -2 and -7
write particle files; any other element type is refused with an error, as are
decks outside Flagbc=1 and Flagdist=19/199.

Output files are those of diagnostic1_Output (src/Contrl/Output.f90:341-352):
fort.18, fort.24-27, fort.29 and fort.30, one row at z = 0 and one after each
space-charge step, values separated by blanks.
"""
from __future__ import annotations

import argparse
import dataclasses
import itertools
import os
from dataclasses import dataclass, field

import numpy as np

from . import hip
from . import particle_io as PIO
from .beam import Beam, Element, rfdata_name
from .sim import Simulation

# Zero-length elements that only write files in the reference
# (src/Contrl/AccSimulator.f90:883-922); they leave the bunch unchanged.
DIAGNOSTIC_TYPES = {-2: "phase-space output", -3: "accumulated density output",
                    -4: "1D density output", -5: "2D density output",
                    -6: "3D density output", -7: "particle dump",
                    -8: "slice output"}


class DeckError(ValueError):
    pass


@dataclass
class DeckElement:
    length: float
    nseg: int
    mapstp: int
    type: int
    values: list[float] = field(default_factory=list)


@dataclass
class Deck:
    npcol: int
    nprow: int
    dim: int
    np: int
    flagmap: int
    flagerr: int
    flagdiag: int
    nx: int
    ny: int
    nz: int
    flagbc: int
    xrad: float
    yrad: float
    perdlen: float
    flagdist: int
    rstartflg: int
    flagsbstp: int
    nchrg: int
    nptlist: list[int]
    currlist: list[float]
    qmcclist: list[float]
    distparam: list[float]
    bcurr: float
    bkenergy: float
    bmass: float
    bcharge: float
    bfreq: float
    phsini: float
    elements: list[DeckElement]

    @property
    def grid(self) -> tuple[int, int, int]:
        return (self.nx, self.ny, self.nz)

    def beam(self) -> Beam:
        return Beam(current=self.bcurr, kinetic=self.bkenergy, mass=self.bmass,
                    charge=self.bcharge, freq=self.bfreq, phase0=self.phsini)


def _number(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def _records(text: str):
    for line in text.splitlines():
        s = line.strip()
        if not s or s.startswith("!"):
            continue
        yield s


def read_deck(path: str) -> Deck:
    """Parse an ImpactZ.in deck (in1_Input/in2_Input, src/Contrl/Input.f90)."""
    with open(path) as f:
        recs = list(_records(f.read()))
    it = iter(recs)

    def nxt(kind: str) -> list[str]:
        try:
            return next(it).split("/")[0].split()
        except StopIteration:
            raise DeckError(f"{path}: deck ends before the {kind} record") from None

    try:
        r = nxt("processor layout")
        npcol, nprow = int(r[0]), int(r[1])
        r = nxt("Dim/Np/Flagmap/Flagerr/Flagdiag")
        dim, npart, flagmap, flagerr, flagdiag = (int(v) for v in r[:5])
        r = nxt("mesh")
        nx, ny, nz, flagbc = (int(v) for v in r[:4])
        xrad, yrad, perdlen = (_number(v) for v in r[4:7])
        r = nxt("distribution flags")
        flagdist, rstartflg, flagsbstp, nchrg = (int(v) for v in r[:4])
        nptlist = [int(v) for v in nxt("particle list")[:nchrg]]
        currlist = [_number(v) for v in nxt("current list")[:nchrg]]
        qmcclist = [_number(v) for v in nxt("q/m list")[:nchrg]]
        distparam = []
        for _ in range(3):
            distparam += [_number(v) for v in nxt("distribution parameters")[:7]]
        r = nxt("beam")
        bcurr, bkenergy, bmass, bcharge, bfreq, phsini = (_number(v) for v in r[:6])
    except (ValueError, IndexError) as e:
        raise DeckError(f"{path}: malformed deck header ({e})") from None
    elements = []
    for line in it:
        toks = line.split("/")[0].split()
        if not toks:
            continue
        try:
            elements.append(DeckElement(_number(toks[0]), int(_number(toks[1])),
                                        int(_number(toks[2])), int(_number(toks[3])),
                                        [_number(v) for v in toks[4:]]))
        except (ValueError, IndexError):
            raise DeckError(f"{path}: malformed lattice line: {line}") from None
    return Deck(npcol, nprow, dim, npart, flagmap, flagerr, flagdiag, nx, ny, nz,
                flagbc, xrad, yrad, perdlen, flagdist, rstartflg, flagsbstp, nchrg,
                nptlist, currlist, qmcclist, distparam, bcurr, bkenergy, bmass,
                bcharge, bfreq, phsini, elements)


@dataclass
class FileAction:
    """A zero-length element that writes a particle file at its place in the
    lattice (src/Contrl/AccSimulator.f90:883-911): -2 phase dump
    (phase_Output, sample period = v1) or -7 checkpoint (outpoint_Output);
    the file is fort.<mapstp>."""
    index: int          # 1-based element index in the deck
    type: int
    unit: int
    value: float = 0.0


@dataclass
class WakeSwitch:
    """A type -41 element: the read-in wakefield on or off
    (src/Contrl/AccSimulator.f90:960-980); its wake functions come from
    rfdataN.in, N = v2."""
    index: int
    element: Element


@dataclass
class MapSwitch:
    """A type -25 element: Flagmap = its mapstp from here on
    (src/Contrl/AccSimulator.f90:807-809)."""
    index: int
    flagmap: int


def program_of(deck: Deck) -> tuple[list, list[str]]:
    """The deck's element table as the sequence this package runs: (deck index,
    Element, FileAction or MapSwitch) in deck order, and notes on the
    diagnostic elements skipped.  Raises DeckError for anything the GPU path
    does not implement under the integrator in effect at that element."""
    def integrator(flagmap, where):
        if flagmap not in (1, 2):
            raise DeckError(f"{where}Flagmap={flagmap}: the linear map (1) and nonlinear "
                            "Lorentz (2) integrators are on this path")
        return flagmap
    flagmap = integrator(deck.flagmap, "")
    if deck.flagbc != 1:
        raise DeckError(f"Flagbc={deck.flagbc}: only open boundaries (Flagbc=1)")
    if deck.flagsbstp != 0:
        raise DeckError("Flagsbstp != 0 (sub-cycling) is not on this path")
    if not 1 <= deck.nchrg <= 100:
        raise DeckError(f"{deck.nchrg} charge states: the reference holds 1 to 100")
    out, notes = [], []
    # whether the run loop has read Fourier coefficients (read3_Data) into
    # the Data module's Fcoef by the time it reaches an element: SolRF under
    # either integrator, every cavity under Flagmap=2
    # (src/Contrl/AccSimulator.f90:818-849)
    fcoef_read = False
    for k, e in enumerate(deck.elements, start=1):
        v = e.values + [0.0] * 4
        if e.type > 100 and (e.type in (105, 106) or flagmap != 1):
            fcoef_read = True
        if e.type == 0:
            out.append((k, Element(e.length, e.nseg, e.mapstp, 0, v[0])))
        elif e.type == 1:
            # map1 reads the field file when its ID > 1e-5; getfld_Quadrupole
            # (Flagmap=2) when it is > 0 (src/Appl/Quadrupole.f90:379-400)
            if v[1] > (1.0e-5 if flagmap == 1 else 0.0):
                raise DeckError(f"element {k}: quadrupole with field file ID {v[1]:g} "
                                "(fringe field from file) is not on this path")
            if flagmap == 1 and v[1] < -10.0:
                raise DeckError(f"element {k}: undulator map (file ID {v[1]:g}) is not "
                                "on this path")
            out.append((k, Element(e.length, e.nseg, e.mapstp, 1, v[0], v[1], v[2], v[3])))
        elif e.type in (2, 3):
            # drift1_BeamBunch takes its ideal-RF branch when Param(5) < 0
            # (src/Appl/BeamBunch.f90:324-330); Flagmap=2 does not call it
            if flagmap == 1 and v[3] < 0.0:
                raise DeckError(f"element {k}: type {e.type} with Param(5) = {v[3]:g} < 0 "
                                "(ideal RF cavity model) is not on this path")
            out.append((k, Element(e.length, e.nseg, e.mapstp, e.type, v[0], v[1], v[2], v[3])))
        elif 101 <= e.type <= 105:
            vv = (v + [0.0] * 14)[:14]
            if flagmap == 1 and vv[3] < 0.0:
                raise DeckError(f"element {k}: type {e.type} with Param(5) = {vv[3]:g} < 0 "
                                "(ideal RF cavity model) is not on this path")
            out.append((k, Element(e.length, e.nseg, e.mapstp, e.type, *vv[:4],
                                   more=tuple(vv[4:]))))
        elif e.type == 4:
            # with current the run loop adds space charge, the CSR wake (v3 >
            # 200 in the bend, v3 > 500 also in every element after it) and
            # ISRchicane, whose deviates come from the compiler's random
            # generator: here from a seeded one (src/Contrl/AccSimulator.f90:
            # 1032-1056, 1131-1138, 1344-1396, 1430-1448); Flagmap=1 only
            if deck.bcurr > 1.0e-15 and flagmap != 1:
                raise DeckError(f"element {k}: a bend (type 4) with current under "
                                "Flagmap=2 is not on this path")
            if deck.bcurr > 1.0e-15:
                notes.append(f"element {k}: bend with current: ISRchicane deviates from a "
                             "seeded generator, not the reference executable's stream")
            vv = (v + [0.0] * 14)[:14]
            if not (e.length > 0.0 and vv[0] != 0.0):
                raise DeckError(f"element {k}: a bend needs a length > 0 and an angle != 0")
            # under Flagmap=2 the loop also kicks a bend with getfld_Dipole,
            # whose pole-face geometry is the Data module's Fcoef
            # (src/Appl/Dipole.f90:361-414): zero, and the kick the identity,
            # until a cavity's Fourier coefficients have been read into it
            if flagmap == 2 and fcoef_read:
                raise DeckError(f"element {k}: a bend under Flagmap=2 after an RF cavity's "
                                "Fourier coefficients were read (getfld_Dipole then reads "
                                "them as its pole-face geometry) is not on this path")
            out.append((k, Element(e.length, e.nseg, e.mapstp, 4, *vv[:4],
                                   more=tuple(vv[4:]))))
        elif e.type == -41:
            vv = (v + [0.0] * 14)[:14]
            out.append((k, WakeSwitch(k, Element(e.length, e.nseg, e.mapstp, -41, *vv[:4],
                                                 more=tuple(vv[4:])))))
        elif e.type == -25:
            if e.nseg != 0:
                raise DeckError(f"element {k}: a type -25 element with {e.nseg} segments "
                                "(zero-length space-charge steps) is not on this path")
            flagmap = integrator(e.mapstp, f"element {k}: ")
            out.append((k, MapSwitch(k, flagmap)))
        elif e.type in DIAGNOSTIC_TYPES and e.length == 0.0 and e.nseg == 0:
            if e.type in (-2, -7):
                out.append((k, FileAction(k, e.type, e.mapstp, v[0])))
            else:
                notes.append(f"element {k}: type {e.type} ({DIAGNOSTIC_TYPES[e.type]}) "
                             "skipped: it writes files only")
        else:
            raise DeckError(f"element {k}: type {e.type} is not on this path "
                            "(drift 0, quadrupole 1, constant focusing 2, solenoid 3 and the "
                            "RF cavities 101-105 are, and bends (4))")
    return out, notes


def lattice_of(deck: Deck) -> tuple[list[Element], list[str]]:
    """The element table of the deck as this package's Elements, and notes on
    the file-writing elements left out of it."""
    prog, notes = program_of(deck)
    out = []
    for _, p in prog:
        if isinstance(p, FileAction):
            notes.append(f"element {p.index}: type {p.type} "
                         f"({DIAGNOSTIC_TYPES[p.type]}) writes fort.{p.unit}")
        elif isinstance(p, MapSwitch):
            notes.append(f"element {p.index}: type -25 switches to Flagmap={p.flagmap}")
        elif isinstance(p, WakeSwitch):
            notes.append(f"element {p.index}: type -41 switches the read-in wakefield")
        else:
            out.append(p)
    return out, notes


def read_particle_in(path: str) -> np.ndarray:
    """particle.in of readin_Dist (src/Appl/Distribution.f90:1677-1705): the
    count, then 9 values per particle (Fortran d exponents accepted)."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        body = f.read()
    if "d" in body or "D" in body:
        body = body.replace("d", "e").replace("D", "e")
    vals = np.array(body.split(), dtype=np.float64) if n else np.zeros(0)
    if vals.size != 9 * n:
        raise DeckError(f"{path}: expected {n} particles of 9 values, got {vals.size} values")
    return vals.reshape(n, 9)


FORT_FILES = (18, 24, 25, 26, 27, 28, 29, 30, 32)


def write_fort(outdir: str, rows: np.ndarray) -> None:
    """The files of diagnostic1_Output (src/Contrl/Output.f90:330-352) from
    diagnostic rows: z, the 41 values of impactz.sim.diag_row, then Nptlocal,
    Npt and the particle count of each charge state.  fort.18/24-27/29/30
    hold the reals; fort.28 z, the smallest and largest local count and Npt
    (one process: both local counts are Nptlocal); fort.32 z and the counts
    per charge state."""
    rows = np.asarray(rows, dtype=np.float64)
    cols = {18: slice(1, 6), 24: slice(6, 12), 25: slice(12, 18), 26: slice(18, 24),
            27: slice(24, 30), 29: slice(30, 36), 30: slice(36, 42)}
    for unit in FORT_FILES:
        with open(os.path.join(outdir, f"fort.{unit}"), "w") as f:
            for r in rows:
                if unit == 28:
                    n = [int(round(v)) for v in r[42:44]]
                    f.write(f"{float(r[0])!r} {n[0]:10d}{n[0]:10d}{n[1]:10d}\n")
                elif unit == 32:
                    f.write(" ".join([repr(float(r[0]))] +
                                     [str(int(round(v))) for v in r[44:]]) + "\n")
                else:
                    vals = [r[0], *r[cols[unit]]]
                    f.write(" ".join(repr(float(x)) for x in vals) + "\n")


def read_rfdata(path: str, per_record: int) -> np.ndarray:
    """An rfdataN.in as read1_Data (per_record = 4: z, Ez, Ez', Ez'') or
    read3_Data (per_record = 1) read it (src/DataStruct/Data.f90:183-192,
    387-392): each list-directed READ takes per_record values, continuing onto
    following lines if needed and dropping the rest of its last line."""
    rows, cur = [], []
    with open(path) as f:
        for line in f:
            toks = line.replace("d", "e").replace("D", "e").split()
            for t in toks:
                cur.append(float(t))
                if len(cur) == per_record:
                    rows.append(cur)
                    cur = []
                    break
    if not rows:
        raise DeckError(f"{path}: no field data")
    a = np.array(rows, dtype=np.float64)
    return a if per_record > 1 else a.ravel()


def with_fields(prog, workdir: str, flagmap: int = 1):
    """The program with each RF cavity's rfdataN.in (N = its file ID v4)
    attached, read once per file and reader as run_AccSimulator does
    (src/Contrl/AccSimulator.f90:818-849): Fourier coefficients (read3_Data)
    for SolRF and for every cavity under Flagmap=2, else the field table
    (read1_Data).  ``flagmap`` is the deck's; -25 switches change it."""
    cache, out = {}, []
    for index, p in prog:
        if isinstance(p, MapSwitch):
            flagmap = p.flagmap
        if isinstance(p, Element) and 101 <= p.type <= 105:
            fid = int(p.v4 + 0.1)
            fourier = p.type == 105 or flagmap != 1
            key = (fid, fourier)
            if key not in cache:
                cache[key] = read_rfdata(os.path.join(workdir, rfdata_name(fid)),
                                         1 if fourier else 4)
            p = dataclasses.replace(p, field=cache[key])
        elif isinstance(p, WakeSwitch) and p.element.v2 > 0.0:
            # read1wk_Data (src/DataStruct/Data.f90:559-633)
            fid = int(p.element.v2 + 0.1)
            key = (fid, False)
            if key not in cache:
                cache[key] = read_rfdata(os.path.join(workdir, rfdata_name(fid)), 4)
            p = WakeSwitch(p.index, dataclasses.replace(p.element, field=cache[key]))
        out.append((index, p))
    return out


def initial_particles(deck: Deck, workdir: str) -> np.ndarray:
    """The initial particles: particle.in (Flagdist=19, readin_Dist) or the
    binary fort.N of readPar_Dist (Flagdist=199, N = distparam(2))."""
    if deck.flagdist == 19:
        return read_particle_in(os.path.join(workdir, "particle.in"))
    if deck.flagdist == 199:
        unit = int(deck.distparam[1] + 0.001)
        return PIO.read_binary_particles(os.path.join(workdir, f"fort.{unit}"),
                                         deck.distparam)
    raise DeckError(f"Flagdist={deck.flagdist}: inputs come from particle.in "
                    "(Flagdist=19) or a binary fort.N (199) on this path")


def run_deck(workdir: str, precision: int = hip.FP64, fused: bool = False,
             outdir: str | None = None, nsteps: int | None = None,
             isr_seed: int = 0) -> np.ndarray:
    """Run ``workdir/ImpactZ.in`` and write the diagnostic files into
    ``outdir`` (default: workdir).  fused=True runs each stretch of
    linear-map (Flagmap=1) segments between file-writing elements through
    izh_run; Flagmap=2 segments run step by step.  Bends with current draw
    ISRchicane's deviates from numpy's generator seeded with isr_seed.
    Rstartflg=1 restarts from the -7 checkpoint fort.<mapstp of the last -7
    element> at the element after the one that wrote it, as
    run_AccSimulator does (src/Contrl/AccSimulator.f90:558-561,674-742).
    Returns the diagnostic rows."""
    deck = read_deck(os.path.join(workdir, "ImpactZ.in"))
    prog, notes = program_of(deck)
    prog = with_fields(prog, workdir, deck.flagmap)
    for n in notes:
        print(f"[impactz.deck] {n}", flush=True)
    out = outdir or workdir
    # counters of run_AccSimulator that the checkpoint records
    st = {"z": 0.0, "i": 0, "j": 0, "ibal": 0, "nstep": 0}
    ck = None
    if deck.rstartflg == 1:
        units = [p.unit for _, p in prog if isinstance(p, FileAction) and p.type == -7]
        if not units:
            raise DeckError("Rstartflg=1 needs a -7 element naming the checkpoint file")
        ck = PIO.read_checkpoint(os.path.join(workdir, f"fort.{units[-1]}"))
        pts = ck.pts
        st.update(z=ck.z, i=ck.inb, j=ck.jstp, ibal=ck.ibal, nstep=ck.nstep)
    else:
        pts = initial_particles(deck, workdir)
    elements = [p for _, p in prog if isinstance(p, Element)]
    sim = Simulation(pts, deck.beam(), elements, deck.grid, precision=precision,
                     flagmap=deck.flagmap, qmcclist=deck.qmcclist, nptlist=deck.nptlist)
    sim.isr_rng = np.random.default_rng(isr_seed)

    def counts():
        info = sim.bunch.info()
        return [info.nptlocal, info.npt, *sim.nptlist]
    try:
        if ck is not None:
            # inpoint_Output: the particles, refptcl and z of the checkpoint;
            # Current stays the deck's and Npt is the particle count
            sim.bunch.set_refpt(ck.refpt)
            sim.z = ck.z
        rows = [np.concatenate([[sim.z], sim.diag(), counts()])]
        sc = deck.bcurr >= float(np.float32(1.0e-10))
        budget = nsteps
        pending = []   # (element, tau1) of segments not yet run

        def flush():
            if not pending:
                return
            # stretches of linear-map steps fuse; bends run step by step
            for bend, grp in itertools.groupby(pending, key=lambda p: p[0].type == 4):
                grp = list(grp)
                # a charge-state count of a bunch of several q/mc^2 scans the
                # particles after each step, which the fused run does not stop for
                if fused and sim.flagmap == 1 and not bend and not sim.chg_needs_scan:
                    z = sim.z
                    recs = sim.run_elements([e for e, _ in grp])
                    for (e, tau1), rec in zip(grp, recs):
                        z = (z + tau1) + tau1  # drift1_BeamBunch advances z by tau1 twice
                        npt = int(rec[44])
                        chg = sim.update_chg(npt)
                        rows.append(np.concatenate([[z], sim.diag(rec), [npt, npt], chg]))
                else:
                    for e, _ in grp:
                        mom = sim.step(e)
                        sim.update_chg()
                        rows.append(np.concatenate([[sim.z], sim.diag(mom), counts()]))
            pending.clear()

        for index, p in prog:
            if index <= st["i"]:
                # a restart resumes after the checkpoint's element, with the
                # deck's Flagmap (src/Contrl/AccSimulator.f90:742)
                continue
            if isinstance(p, MapSwitch):
                flush()
                sim.flagmap = p.flagmap
                st["i"], st["j"] = index, 1
                continue
            if isinstance(p, WakeSwitch):
                flush()
                sim.wake_switch(p.element)
                st["i"], st["j"] = index, 1
                continue
            if isinstance(p, FileAction):
                flush()
                _file_action(p, sim, out, st, deck)
                st["i"], st["j"] = index, 1   # its empty segment loop leaves j = 1
                continue
            for _ in range(p.nseg):
                if budget is not None and budget <= 0:
                    break
                pending.append((p, 0.5 * p.length / p.nseg))
                st["nstep"] += 1
                st["ibal"] += 1 if sc else 0
                if budget is not None:
                    budget -= 1
            st["i"], st["j"] = index, (p.nseg + 1 if p.nseg >= 1 else 1)
        flush()
        sim.finish_element()
        rows = np.array(rows)
    finally:
        sim.close()
    write_fort(out, rows)
    return rows


def _file_action(p: FileAction, sim: Simulation, outdir: str, st: dict, deck: Deck):
    sim.finish_element()
    pts = sim.bunch.getpts()
    refpt = sim.bunch.refpt
    if p.type == -2:
        # integerSamplePeriod = realSamplePeriod (conversion truncates)
        PIO.write_phase(os.path.join(outdir, f"fort.{p.unit}"), pts, int(p.value),
                        sim.const.scxl, float(refpt[5]))
    else:
        rng, ms = sim.dom.range, sim.dom.msize
        nx, ny, nz = deck.grid
        ck = PIO.Checkpoint(z=sim.z, inb=p.index, jstp=st["j"], ibal=st["ibal"],
                            nstep=st["nstep"], msize=ms, range=rng,
                            lctabnm=np.array([nz, ny]),
                            lctabrg=np.array([rng[4], rng[5], rng[2], rng[3]]),
                            refpt=refpt, pts=pts)
        PIO.write_checkpoint(os.path.join(outdir, f"fort.{p.unit}"), ck)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("workdir", help="directory holding ImpactZ.in and particle.in")
    ap.add_argument("--mixed", action="store_true", help="mixed-precision form")
    ap.add_argument("--fused", action="store_true", help="run through izh_run")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    hip.init(args.device)
    run_deck(args.workdir, hip.MIXED if args.mixed else hip.FP64, args.fused)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
