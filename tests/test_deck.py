"""The deck path on the CPU: parsing ImpactZ.in records the way in1_Input /
in2_Input read them (src/Contrl/Input.f90:37-384), the element table of this
build, and the refusals for what is not on the path."""
import os

import numpy as np
import pytest

from conftest import fixture_beam, fixture_lattice, load_golden


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


HEADER = """!
! comment lines start with '!'
1 1
!  Dim Np Flagmap Flagerr Flagdiag
6 3 1 0 1
  16 16 32 1 0.02 0.03 0.1
19 0 0 1
3
0.1
1.0657889726d-09
0.001 0.001 0.0 1.0 1.0 0.0 0.0
0.001 0.001 0.0 1.0 1.0 0.0 0.0
0.1 0.001 0.0 1.0 1.0 0.0 0.0
0.1 1.5d8 938.272088d6 1.0 324.0e6 0.0
"""


def test_deck_records_and_lattice(tmp_path):
    from impactz.deck import lattice_of, read_deck
    deck = HEADER + """0.0 0 201 -8 101.0 / slice output
0.4 4 10 0 1.0 /
0.1 4 10 1 10.0 -1.0 0.15 0.0 0.0 0.0 0.0 0.0 /
! a comment inside the lattice
0.0 0 1001 -2 1.0 /
0.4 2 10 0 0.5 /
"""
    p = tmp_path / "ImpactZ.in"
    _write(p, deck)
    d = read_deck(str(p))
    assert (d.npcol, d.nprow, d.dim, d.np, d.flagmap, d.flagerr, d.flagdiag) == (1, 1, 6, 3, 1, 0, 1)
    assert d.grid == (16, 16, 32) and d.flagbc == 1
    assert (d.xrad, d.yrad, d.perdlen) == (0.02, 0.03, 0.1)
    assert (d.flagdist, d.rstartflg, d.flagsbstp, d.nchrg) == (19, 0, 0, 1)
    assert d.qmcclist == [1.0657889726e-09]
    assert len(d.distparam) == 21
    assert (d.bcurr, d.bkenergy, d.bmass, d.bcharge, d.bfreq, d.phsini) == \
        (0.1, 1.5e8, 938.272088e6, 1.0, 324.0e6, 0.0)
    assert [e.type for e in d.elements] == [-8, 0, 1, -2, 0]
    lat, notes = lattice_of(d)
    assert len(notes) == 2
    assert [(e.length, e.nseg, e.mapstp, e.type) for e in lat] == \
        [(0.4, 4, 10, 0), (0.1, 4, 10, 1), (0.4, 2, 10, 0)]
    assert lat[0].radius == 1.0 and lat[2].radius == 0.5
    assert lat[1].v1 == 10.0 and lat[1].radius == 0.15


def test_deck_written_for_the_reference_reads_back(tmp_path):
    """The deck the fixtures were run with (impactz.beam.write_deck) parses
    to the same beam and lattice."""
    from impactz.beam import write_deck
    from impactz.deck import lattice_of, read_deck
    d = load_golden("fodo_traj")
    bm, lat = fixture_beam(d), fixture_lattice(d)
    grid = tuple(int(v) for v in d["grid"])
    p = tmp_path / "ImpactZ.in"
    write_deck(str(p), d["pts_in"].shape[0], grid, bm, lat)
    deck = read_deck(str(p))
    assert deck.grid == grid
    b2 = deck.beam()
    assert (b2.current, b2.kinetic, b2.mass, b2.charge, b2.freq, b2.phase0) == \
        (bm.current, bm.kinetic, bm.mass, bm.charge, bm.freq, bm.phase0)
    lat2, notes = lattice_of(deck)
    assert not notes
    assert [(e.length, e.nseg, e.mapstp, e.type, e.v1, e.radius) for e in lat2] == \
        [(e.length, e.nseg, e.mapstp, e.type, e.v1, e.radius) for e in lat]


@pytest.mark.parametrize("line, what", [
    ("0.1 4 10 106 2503370.0 3.24e8 222 852.0 0.0065 /", "type 106"),
    ("0.1 4 10 1 10.0 2.0 0.15 /", "field file ID"),
    ("0.1 4 10 1 10.0 -20.0 0.15 /", "undulator"),
    ("0.1 4 10 2 30.0 20.0 5.0 -1.0 /", "ideal RF"),
    ("0.0 0 10 -1 1.0 /", "type -1"),
    ("0.0 0 2 -25 /\n0.1 4 10 4 0.02 0.0 150 0.03 /", "with current under Flagmap=2"),
])
def test_deck_refuses_elements_off_the_path(tmp_path, line, what):
    from impactz.deck import DeckError, lattice_of, read_deck
    p = tmp_path / "ImpactZ.in"
    _write(p, HEADER + line + "\n")
    with pytest.raises(DeckError, match=what):
        lattice_of(read_deck(str(p)))


def test_deck_refuses_other_integrators_and_boundaries(tmp_path):
    from impactz.deck import DeckError, lattice_of, read_deck
    for old, new, what in [("6 3 1 0 1", "6 3 3 0 1", "Flagmap=3"),
                           ("16 16 32 1 0.02", "16 16 32 2 0.02", "Flagbc=2")]:
        p = tmp_path / "ImpactZ.in"
        _write(p, HEADER.replace(old, new) + "0.4 4 10 0 1.0 /\n")
        with pytest.raises(DeckError, match=what):
            lattice_of(read_deck(str(p)))


def test_particle_in_round_trip_and_fortran_exponents(tmp_path):
    from impactz.beam import gaussian_bunch, write_particle_in
    from impactz.deck import DeckError, read_particle_in
    pts = gaussian_bunch(50, seed=2)
    p = tmp_path / "particle.in"
    write_particle_in(str(p), pts)
    assert np.array_equal(read_particle_in(str(p)), pts)
    _write(p, "1\n1.0D-3 2.0d0 3 4 5 6 7 8 9\n")
    assert read_particle_in(str(p))[0, 0] == 1.0e-3
    _write(p, "2\n1 2 3 4 5 6 7 8 9\n")
    with pytest.raises(DeckError):
        read_particle_in(str(p))


def test_fort_writer_columns(tmp_path):
    from impactz.deck import write_fort
    rows = np.arange(2 * 46, dtype=np.float64).reshape(2, 46)
    write_fort(str(tmp_path), rows)
    f24 = np.loadtxt(tmp_path / "fort.24")
    assert f24.shape == (2, 7)
    assert np.array_equal(f24[:, 0], rows[:, 0]) and np.array_equal(f24[:, 1:], rows[:, 6:12])
    f18 = np.loadtxt(tmp_path / "fort.18")
    assert np.array_equal(f18[:, 1:], rows[:, 1:6])
    f30 = np.loadtxt(tmp_path / "fort.30")
    assert np.array_equal(f30[:, 1:], rows[:, 36:42])
    # fort.28: z, then Nptlocal twice (smallest and largest over one process)
    # and Npt in I10 fields (src/Contrl/Output.f90:101,348)
    lines = (tmp_path / "fort.28").read_text().splitlines()
    z, ints = lines[1].split(" ", 1)
    assert float(z) == rows[1, 0] and len(ints) == 30
    assert [int(ints[10 * i:10 * i + 10]) for i in range(3)] == [88, 88, 89]
    # fort.32: z and the count of each charge state
    f32 = np.loadtxt(tmp_path / "fort.32")
    assert np.array_equal(f32[:, 1:], rows[:, 44:46]) and np.array_equal(f32[:, 0], rows[:, 0])


def test_deck_reads_charge_states(tmp_path):
    """nchrg, nptlist, currlist and qmcclist as the deck writer writes them
    and Input.f90:106-136 reads them."""
    from impactz.beam import Beam, fodo, write_deck
    from impactz.deck import program_of, read_deck
    p = tmp_path / "ImpactZ.in"
    q = 1.0 / 938.272088e6
    write_deck(str(p), 300, (8, 8, 8), Beam(), fodo(1),
               charge_states=[(100, 0.03, q), (200, 0.07, 0.5 * q)])
    deck = read_deck(str(p))
    assert deck.nchrg == 2 and deck.nptlist == [100, 200]
    assert deck.currlist == [0.03, 0.07] and deck.qmcclist == [q, 0.5 * q]
    prog, _ = program_of(deck)
    assert len(prog) == 4


def test_deck_linear_elements(tmp_path):
    """Constant focusing, solenoid and K-mode quadrupole lines become
    elements with their four values and the reference's pipe radius."""
    from impactz.deck import lattice_of, read_deck
    p = tmp_path / "ImpactZ.in"
    _write(p, HEADER + "0.1 4 10 2 30.0 20.0 5.0 0.02 /\n"
               "0.1 4 10 3 1.5 0.0 0.03 0.0 0.0 /\n"
               "0.1 4 10 1 5.4 -1.0 0.04 /\n")
    lat, notes = lattice_of(read_deck(str(p)))
    assert [e.type for e in lat] == [2, 3, 1]
    assert lat[0].values[:4] == [30.0, 20.0, 5.0, 0.02] and lat[0].radius == 0.02
    assert lat[1].values[:3] == [1.5, 0.0, 0.03] and lat[1].radius == 0.03
    assert lat[2].values[:3] == [5.4, -1.0, 0.04] and lat[2].radius == 0.04
    assert notes == []


def test_deck_integrator_switch(tmp_path):
    """Flagmap=2 decks, the type -25 switch, and the per-integrator refusals:
    a quadrupole field file ID in (0, 1e-5] and ideal-RF Param(5) < 0 are
    refused or accepted according to the integrator in effect."""
    from impactz.deck import DeckError, MapSwitch, program_of, read_deck
    p = tmp_path / "ImpactZ.in"
    _write(p, HEADER.replace("6 3 1 0 1", "6 3 2 0 1") + """0.4 4 10 0 1.0 /
0.1 4 10 1 10.0 -20.0 0.15 /
0.0 0 1 -25 /
0.1 4 10 2 1.0 2.0 3.0 1.0 /
""")
    prog, _ = program_of(read_deck(str(p)))
    assert [type(x).__name__ for _, x in prog] == ["Element", "Element", "MapSwitch",
                                                  "Element"]
    assert prog[2][1] == MapSwitch(3, 1)
    _write(p, HEADER.replace("6 3 1 0 1", "6 3 2 0 1") + "0.1 4 10 1 10.0 1e-6 0.15 /\n")
    with pytest.raises(DeckError, match="file ID"):
        program_of(read_deck(str(p)))
    _write(p, HEADER + "0.0 0 2 -25 /\n0.1 4 10 2 1.0 2.0 3.0 -1.0 /\n0.0 0 1 -25 /\n"
           "0.1 4 10 2 1.0 2.0 3.0 -1.0 /\n")
    with pytest.raises(DeckError, match="element 4: type 2 .*ideal RF"):
        program_of(read_deck(str(p)))
    for line, what in (("0.0 0 3 -25 /", "element 1: Flagmap=3"),
                       ("0.0 2 2 -25 /", "segments")):
        _write(p, HEADER + line + "\n")
        with pytest.raises(DeckError, match=what):
            program_of(read_deck(str(p)))


def test_deck_rf_cavities_and_field_files(tmp_path):
    """RF cavity lines carry v1..v14 and the pipe radius v5; their rfdataN.in
    is read the way read1_Data / read3_Data read it (one list-directed READ
    per row: 4 values for a table, 1 for Fourier coefficients, continuing
    onto the next line when short and dropping the rest of a long one)."""
    from impactz.deck import (DeckError, program_of, read_deck, read_rfdata, with_fields)
    p = tmp_path / "ImpactZ.in"
    _write(p, HEADER + """0.2345 6 20 101 5.0d6 324.0e6 -30.0 1.0 0.02 0.03 12.0 0.03 -12.0 /
0.2 6 20 105 3.0e6 324.0e6 45.0 2.0 0.015 0 0 0 0 0 0.4 /
""")
    _write(tmp_path / "rfdata1.in", "0.1 1.0 0.0 0.0 extra\n0.2 2.0\n0.5 0.0\n0.3 3.0 1.0 2.0\n")
    _write(tmp_path / "rfdata2.in", "7.0 ignored\n0.0\n0.2\n0.2\n0.4\n-1.0\n0.05\n")
    prog, _ = program_of(read_deck(str(p)))
    dtl, slrf = prog[0][1], prog[1][1]
    assert dtl.type == 101 and dtl.values[:9] == [5e6, 324e6, -30.0, 1.0, 0.02, 0.03, 12.0,
                                                  0.03, -12.0]
    assert dtl.radius == 0.02 and slrf.radius == 0.015 and slrf.values[10] == 0.4
    tab = read_rfdata(str(tmp_path / "rfdata1.in"), 4)
    assert tab.tolist() == [[0.1, 1.0, 0.0, 0.0], [0.2, 2.0, 0.5, 0.0], [0.3, 3.0, 1.0, 2.0]]
    coef = read_rfdata(str(tmp_path / "rfdata2.in"), 1)
    assert coef.tolist() == [7.0, 0.0, 0.2, 0.2, 0.4, -1.0, 0.05]
    withf = with_fields(prog, str(tmp_path))
    assert np.array_equal(withf[0][1].field, tab) and np.array_equal(withf[1][1].field, coef)
    for line, what in (("0.2 4 20 103 1e6 324e6 0.0 -1.0 0.02 /", "ideal RF"),
                       ("0.2 4 20 106 1e6 324e6 0.0 2.0 0.02 /", "type 106")):
        _write(p, HEADER + line + "\n")
        with pytest.raises(DeckError, match=what):
            program_of(read_deck(str(p)))
    # under Flagmap=2 the DTL family reads Fourier coefficients (read3_Data)
    _write(p, HEADER.replace("6 3 1 0 1", "6 3 2 0 1") +
           "0.2 4 20 103 1e6 324e6 0 1 0.02 /\n")
    _write(tmp_path / "rfdata1.in", "0.5\n1.0\n0.0\n")
    prog = with_fields(program_of(read_deck(str(p)))[0], str(tmp_path), 2)
    assert prog[0][1].field.tolist() == [0.5, 1.0, 0.0]


def test_deck_wake_switches(tmp_path):
    """-41 lines become wake switches carrying their wake-function file
    (rfdataN.in, N = v2); SolRF lines keep their wake values v12..v14."""
    from impactz.deck import DeckError, WakeSwitch, program_of, read_deck, with_fields
    p = tmp_path / "ImpactZ.in"
    _write(p, HEADER + """0.0 0 1 -41 1.5 3.0 20.0 /
0.4 4 10 0 1.0 /
0.2 6 20 105 3.0e6 324.0e6 45.0 2.0 0.015 0 0 0 0 0 0.4 10.75d-3 30.33d-3 33.321d-3 /
""")
    _write(tmp_path / "rfdata3.in", "0.0 1e14 1e18 2e18\n0.001 2e14 3e18 4e18\n")
    _write(tmp_path / "rfdata2.in", "3.0\n0.0\n0.2\n0.2\n0.4\n-1.0\n0.05\n")
    prog = with_fields(program_of(read_deck(str(p)))[0], str(tmp_path))
    sw = prog[0][1]
    assert isinstance(sw, WakeSwitch) and sw.element.values[:3] == [1.5, 3.0, 20.0]
    assert sw.element.field.shape == (2, 4) and sw.element.field[1, 2] == 3e18
    assert prog[2][1].values[11:14] == [10.75e-3, 30.33e-3, 33.321e-3]
    # under Flagmap=2 the wake is applied by kick2wake
    _write(p, HEADER.replace("6 3 1 0 1", "6 3 2 0 1") + "0.0 0 1 -41 1.5 3.0 20.0 /\n")
    assert isinstance(program_of(read_deck(str(p)))[0][0][1], WakeSwitch)


def test_deck_bends_with_and_without_current(tmp_path):
    """Bends are taken at zero current and, under Flagmap=1, with current
    (space charge, CSR, ISRchicane with seeded deviates, which the deck's
    notes say); the bend's values reach the element in the deck's order."""
    from impactz import beam as B
    from impactz.deck import program_of, read_deck
    bend = B.Element(0.8, 4, 1, 4, 0.25, 0.5, 0.0, 0.02, more=(0.1, 0.15, 0.5, -0.3, 0.4))
    for cur in (0.0, 0.1):
        path = str(tmp_path / f"deck{cur}.in")
        B.write_deck(path, 10, (8, 8, 8), B.Beam(current=cur), [B.Element(0.3, 2, 1, 0, 1.0),
                                                                 bend])
        prog, notes = program_of(read_deck(path))
        e = prog[1][1]
        assert e.type == 4 and e.values[:9] == pytest.approx(
            [0.25, 0.5, 0.0, 0.02, 0.1, 0.15, 0.5, -0.3, 0.4])
        assert any("seeded generator" in n for n in notes) == (cur > 0)


def test_deck_refuses_lorentz_bend_after_fourier_coefficients(tmp_path):
    """A Flagmap=2 bend runs as its two Sector_Dipole halves because
    getfld_Dipole's geometry (the Data module's Fcoef) is zero; once a
    cavity's Fourier coefficients have been read, it is not, and the deck
    is refused.  The same bend before the cavity, or under Flagmap=1, runs."""
    from impactz.deck import DeckError, program_of, read_deck
    p = tmp_path / "ImpactZ.in"
    hdr = HEADER.replace("0.1 1.5d8 938.272088d6 1.0 324.0e6 0.0",
                         "0.0 1.5d8 938.272088d6 1.0 324.0e6 0.0")
    bend = "0.5 4 1 4 0.1 0.0 0.0 0.02 /\n"
    cav = "0.2 4 20 101 1.0e6 324.0e6 0.0 1.0 1.0 /\n"
    for flagmap, lines, refused in ((2, cav + bend, True), (2, bend + cav, False),
                                   (1, cav + bend, False)):
        _write(p, hdr.replace("6 3 1 0 1", f"6 3 {flagmap} 0 1") + lines)
        deck = read_deck(str(p))
        if refused:
            with pytest.raises(DeckError, match="Fourier coefficients"):
                program_of(deck)
        else:
            program_of(deck)


def test_simulation_refuses_lorentz_bend_after_cavity():
    """The same refusal for a Simulation driven element by element: once a
    cavity's Fourier table has been built under Flagmap=2, a bend raises
    before it touches the bunch."""
    from impactz import beam as B, hip
    from impactz.sim import Simulation
    sim = Simulation.__new__(Simulation)
    sim.beam, sim.flagmap, sim._fcoef = B.Beam(current=0.0), 2, True
    with pytest.raises(hip.IZHError, match="Fourier coefficients"):
        sim.bend_step(None)
    sim.flagmap = 1
    with pytest.raises(AttributeError):   # past the check, into the bunch
        sim.bend_step(None)
