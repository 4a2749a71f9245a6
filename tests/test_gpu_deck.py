"""The deck path on the GPU: an ImpactZ.in + particle.in run through
impactz.deck must write the diagnostic files of the reference (every column
of fort.18, 24-27, 29 and 30 within 1e-10 relative, fort.28 and fort.32
exactly), step by step and through izh_run."""
import os

import numpy as np
import pytest

from conftest import deck_files_err, fixture_beam, fixture_lattice, load_golden
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,fused", [("fodo_traj", False), ("fodo_traj", True),
                                        ("kquad", False), ("cf_sol", True),
                                        ("nl_fodo", True), ("nl_cf_sol", False),
                                        ("rf", True), ("rf", False)])
def test_deck_run_matches_reference_fort_files(tmp_path, gpu, name, fused):
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    flagmap = int(d["flagmap"]) if "flagmap" in d else 1
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid,
               fixture_beam(d), fixture_lattice(d), flagmap=flagmap)
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    nsteps = int(d["nsteps"])
    print(f"deck run {name} fused={fused} flagmap={flagmap}", flush=True)
    run_deck(str(tmp_path), fused=fused, nsteps=nsteps)
    assert d["fort24"].shape[0] == nsteps + 1
    worst = deck_files_err(str(tmp_path), d)
    record(test="deck_run", case=name, fused=fused, steps=nsteps, moments=worst)
    assert worst < TOL


@pytest.mark.parametrize("tag,fused", [("sw", True), ("nl", False)])
def test_deck_lorentz_runs_match_reference_executable(tmp_path, gpu, tag, fused):
    """Decks the reference executable ran (tests/golden/exe_nonlinear.npz):
    Flagmap=1 switched to 2 by a -25 element (fused up to the switch), and a
    Flagmap=2 FODO period."""
    from impactz.beam import Element, write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden("exe_nonlinear")
    grid = tuple(int(v) for v in d["grid"])
    lat = [Element(float(r[0]), int(r[1]), int(r[2]), int(r[3]),
                   *(float(v) for v in r[4:8])) for r in d[f"{tag}_lattice"]]
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid, fixture_beam(d),
               lat, flagmap=int(d[f"{tag}_flagmap"]))
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    print(f"deck run exe_nonlinear:{tag} fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    worst = deck_files_err(str(tmp_path), d, f"{tag}_")
    record(test="deck_exe_lorentz", case=tag, fused=fused, moments=worst)
    assert worst < TOL


@pytest.mark.parametrize("fused", [True, False])
def test_deck_charge_states_match_reference_executable(tmp_path, gpu, fused):
    """Two charge states through a lossy FODO as the reference executable ran
    them (tests/golden/exe_chg.npz): the q/mc^2 of the particles vary, so
    the deck runs step by step and recounts the states on the device after
    every step (izh_chgupdate); fort.28 and fort.32 must equal the
    executable's, every other column within 1e-10."""
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden("exe_chg")
    grid = tuple(int(v) for v in d["grid"])
    cs = [(int(r[0]), float(r[1]), float(r[2])) for r in d["charge_states"]]
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid, fixture_beam(d),
               fixture_lattice(d), charge_states=cs)
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    print(f"deck run exe_chg fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    worst = deck_files_err(str(tmp_path), d)
    record(test="deck_charge_states", fused=fused, moments=worst)
    assert worst < TOL


def test_chgupdate_matches_oracle(gpu, oracle):
    """izh_chgupdate against the oracle's chgupdate_BeamBunch on a bunch of
    three q/mc^2 (one a hair inside the 1e-4 tolerance of another), after
    particles were lost, for a list in which some particles match two
    states."""
    from impactz import beam as B
    from impactz import hip
    bm = B.Beam()
    pts = B.gaussian_bunch(5000, seed=21, beam=bm)
    q = bm.charge / bm.mass
    eps = float(np.float32(1.0e-4))
    pts[:, 6] = np.where(np.arange(5000) % 3 == 0, q,
                         np.where(np.arange(5000) % 3 == 1, 0.5 * q, q * (1 + 0.6 * eps)))
    b = hip.BeamBunch.from_beam(pts, bm)
    try:
        b.conv1st(0.05, 0.0015, 0.0015)      # loses the outer particles
        b.cvbkforth1st()
        live = b.getpts()
        assert 0 < live.shape[0] < 5000
        for qlist in ([q], [q, 0.5 * q], [q * (1 + 0.3 * eps), 0.5 * q, 2 * q]):
            got = b.chgupdate(qlist)
            want = oracle.chgupdate(live, qlist)
            assert np.array_equal(got, want), (qlist, got, want)
    finally:
        b.close()
