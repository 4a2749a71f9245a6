"""The deck path on the GPU: an ImpactZ.in + particle.in run through
impactz.deck must write fort.24/25/26 rows equal to the reference's (1e-10
relative), step by step and through izh_run."""
import os

import numpy as np
import pytest

from conftest import diag_row_err, fixture_beam, fixture_lattice, load_golden
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
    f = {u: np.loadtxt(tmp_path / f"fort.{u}", ndmin=2) for u in (18, 24, 25, 26)}
    assert f[24].shape == (nsteps + 1, 7)
    worst = 0.0
    for k in range(1, nsteps + 1):
        got = np.concatenate([[0.0] * 5, f[24][k, 1:], f[25][k, 1:], f[26][k, 1:]])
        ref = np.concatenate([d["fort24"][k, 1:], d["fort25"][k, 1:], d["fort26"][k, 1:]])
        worst = max(worst, diag_row_err(got, ref, ref_is_fort=True))
    assert np.max(np.abs(f[24][:, 0] - d["fort24"][: nsteps + 1, 0])) < 1e-12
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
    f = {u: np.loadtxt(tmp_path / f"fort.{u}", ndmin=2) for u in (24, 25, 26)}
    ref24 = d[f"{tag}_fort24"]
    assert f[24].shape == ref24.shape
    worst = 0.0
    for k in range(1, ref24.shape[0]):
        got = np.concatenate([[0.0] * 5, f[24][k, 1:], f[25][k, 1:], f[26][k, 1:]])
        ref = np.concatenate([ref24[k, 1:], d[f"{tag}_fort25"][k, 1:],
                              d[f"{tag}_fort26"][k, 1:]])
        worst = max(worst, diag_row_err(got, ref, ref_is_fort=True))
    assert np.max(np.abs(f[24][:, 0] - ref24[:, 0])) < 1e-12
    record(test="deck_exe_lorentz", case=tag, fused=fused, moments=worst)
    assert worst < TOL
