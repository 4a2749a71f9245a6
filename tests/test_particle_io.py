"""Particle files of the reference (impactz.particle_io) against files the
compiled reference wrote itself (tests/golden/io_ckpt.npz): its -7
checkpoint, its -2 phase dump, and its run on a binary fort.40 written here
(Flagdist=199)."""
import numpy as np
import pytest

from conftest import fixture_beam, load_golden, rel_err


def _ref_ckpt(tmp_path):
    d = load_golden("io_ckpt")
    p = tmp_path / "fort.50"
    p.write_bytes(d["ckpt50"].tobytes())
    return d, str(p)


def test_checkpoint_reader_on_the_reference_file(tmp_path):
    from impactz import particle_io as P
    d, path = _ref_ckpt(tmp_path)
    ck = P.read_checkpoint(path)
    # outpoint_Output at the -7 element (deck index 4) after 8 steps: j is 1
    # (the empty loop of the -2 element before it), 8 space-charge steps
    assert (ck.inb, ck.jstp, ck.ibal, ck.nstep) == (4, 1, 8, 8)
    assert ck.z == pytest.approx(d["a_fort18"][8, 0], abs=1e-15)
    assert ck.refpt[4] == pytest.approx(d["a_fort18"][8, 1], rel=1e-12)
    assert ck.refpt[5] == pytest.approx(-d["a_fort18"][8, 2], rel=1e-12)
    assert list(ck.lctabnm) == [16, 16]
    assert ck.pts.shape == (1000, 9)
    # the -2 dump the reference wrote at the same place holds the same particles
    assert np.array_equal(ck.pts, d["phase60"])
    # the domain of the last update: h = extent / (N - 1)
    assert np.allclose(ck.msize * (np.array([16, 16, 16]) - 1),
                       ck.range[1::2] - ck.range[0::2], rtol=1e-12)


def test_checkpoint_writer_reproduces_the_reference_bytes(tmp_path):
    from impactz import particle_io as P
    d, path = _ref_ckpt(tmp_path)
    ck = P.read_checkpoint(path)
    out = tmp_path / "again"
    P.write_checkpoint(str(out), ck)
    assert out.read_bytes() == d["ckpt50"].tobytes()


def test_fortran_records_are_checked(tmp_path):
    from impactz import particle_io as P
    d, path = _ref_ckpt(tmp_path)
    raw = d["ckpt50"].tobytes()
    bad = tmp_path / "bad"
    bad.write_bytes(raw[:-5])
    with pytest.raises(P.FortranFileError):
        P.read_checkpoint(str(bad))
    bad.write_bytes(raw[:12])
    with pytest.raises(P.FortranFileError):
        P.read_checkpoint(str(bad))


def test_binary_particles_as_the_reference_read_them(tmp_path, oracle):
    """readPar_Dist: the reference ran on a fort.40 written by
    write_binary_particles; its initial fort.24-26 row is that of the
    particles plus the distparam offsets."""
    from impactz import particle_io as P
    d = load_golden("io_ckpt")
    p = tmp_path / "fort.40"
    P.write_binary_particles(str(p), d["pts_in"])
    pts = P.read_binary_particles(str(p), d["c_distparam"])
    dp = d["c_distparam"]
    shift = np.array([dp[5], dp[6], dp[12], dp[13], dp[19], dp[20]])
    assert np.array_equal(pts[:, :6], d["pts_in"][:, :6] + shift)
    assert np.array_equal(pts[:, 6:], d["pts_in"][:, 6:])
    bm = fixture_beam(d)
    c = oracle.physconst(bm.freq)
    s, m, r2 = oracle.diag_sums(pts)
    refpt = np.zeros(6)
    refpt[4], refpt[5] = bm.phase0, -bm.gamma0
    row = oracle.diag_row(c, s, m, r2, pts.shape[0], refpt, bm.mass)
    ref = np.concatenate([d["c_fort24"][0, 1:], d["c_fort25"][0, 1:], d["c_fort26"][0, 1:]])
    assert rel_err(row[5:23], ref) < 1e-12


def test_phase_writer(tmp_path):
    from impactz import particle_io as P
    d = load_golden("io_ckpt")
    pts = d["phase60"]
    f = tmp_path / "fort.60"
    P.write_phase(str(f), pts, 1, 0.147, -1.16)
    assert np.array_equal(P.read_phase(str(f)), pts)
    P.write_phase(str(f), pts, 3, 0.147, -1.16)
    assert np.array_equal(P.read_phase(str(f)), pts[::3])
    # samplePeriod < 0: the six Impact-T columns
    P.write_phase(str(f), pts, -2, 0.147, -1.16)
    got = P.read_phase(str(f))
    sel = pts[::2]
    want = np.stack([sel[:, 0] * 0.147, sel[:, 1], sel[:, 2] * 0.147, sel[:, 3],
                     -sel[:, 4] * 0.147, -sel[:, 5] + 1.16], axis=1)
    assert np.array_equal(got, want)


def test_deck_program_has_file_actions(tmp_path):
    from conftest import fixture_lattice
    from impactz import beam as B
    from impactz.deck import FileAction, lattice_of, program_of, read_deck
    d = load_golden("io_ckpt")
    lat = fixture_lattice(d)
    B.write_deck(str(tmp_path / "ImpactZ.in"), 1000, (16, 16, 16), fixture_beam(d), lat)
    prog, notes = program_of(read_deck(str(tmp_path / "ImpactZ.in")))
    kinds = [(i, type(p).__name__) for i, p in prog]
    assert kinds == [(1, "Element"), (2, "Element"), (3, "FileAction"), (4, "FileAction"),
                     (5, "Element"), (6, "Element")]
    assert prog[2][1] == FileAction(3, -2, 60, 1.0)
    assert prog[3][1] == FileAction(4, -7, 50, 0.0)
    lat2, notes2 = lattice_of(read_deck(str(tmp_path / "ImpactZ.in")))
    assert len(lat2) == 4 and any("fort.50" in n for n in notes2)
