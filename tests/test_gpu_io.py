"""Particle files on the GPU path: a deck with a -2 phase dump and a -7
checkpoint, its restart (Rstartflg=1) from the checkpoint the reference
wrote, and a run on binary input (Flagdist=199), each against what the
compiled reference wrote for the same decks (tests/golden/io_ckpt.npz)."""
import numpy as np
import pytest

from conftest import (col_err, fixture_beam, fixture_lattice, fixture_row_err, load_golden,
                      rel_err)
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu


def _rows_err(tmp_path, d, prefix, nrows):
    """Every column of the diagnostic files the deck wrote against the
    executable's; fort.28 and fort.32 exactly."""
    from conftest import FORT_COLS
    f = {u: np.loadtxt(tmp_path / f"fort.{u}", ndmin=2) for u, _, _ in FORT_COLS}
    assert f[24].shape == (nrows, 7)
    worst = 0.0
    for k in range(nrows):
        got = np.zeros(41)
        for u, lo, n in FORT_COLS:
            got[lo:lo + n] = f[u][k, 1:1 + n]
        worst = max(worst, fixture_row_err(got, d, k, f"{prefix}_"))
    assert np.max(np.abs(f[24][:, 0] - d[f"{prefix}_fort24"][:nrows, 0])) < 1e-12
    for u in (28, 32):
        mine = np.loadtxt(tmp_path / f"fort.{u}", ndmin=2)
        assert np.array_equal(mine[:, 1:], d[f"{prefix}_fort{u}"][:nrows, 1:]), u
    return worst


@pytest.mark.parametrize("fused", [False, True])
def test_deck_phase_dump_and_checkpoint(tmp_path, gpu, fused):
    from impactz import particle_io as P
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden("io_ckpt")
    grid = tuple(int(v) for v in d["grid"])
    write_deck(str(tmp_path / "ImpactZ.in"), 1000, grid, fixture_beam(d), fixture_lattice(d))
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    print(f"io deck fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    worst = _rows_err(tmp_path, d, "a", 17)
    ours = P.read_checkpoint(str(tmp_path / "fort.50"))
    (tmp_path / "ref.50").write_bytes(d["ckpt50"].tobytes())
    ref = P.read_checkpoint(str(tmp_path / "ref.50"))
    assert (ours.inb, ours.jstp, ours.ibal, ours.nstep) == (ref.inb, ref.jstp, ref.ibal,
                                                             ref.nstep)
    assert ours.z == pytest.approx(ref.z, abs=1e-15)
    err = {"rows": worst, "ckpt_pts": col_err(ours.pts, ref.pts),
           "ckpt_range": rel_err(ours.range, ref.range),
           "ckpt_msize": rel_err(ours.msize, ref.msize),
           "ckpt_refpt": rel_err(ours.refpt, ref.refpt),
           "phase": col_err(P.read_phase(str(tmp_path / "fort.60")), d["phase60"])}
    assert list(ours.lctabnm) == list(ref.lctabnm)
    record(test="deck_io", fused=fused, **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("fused", [False, True])
def test_deck_restart_from_reference_checkpoint(tmp_path, gpu, fused):
    from impactz.beam import write_deck
    from impactz.deck import run_deck
    d = load_golden("io_ckpt")
    grid = tuple(int(v) for v in d["grid"])
    write_deck(str(tmp_path / "ImpactZ.in"), 1000, grid, fixture_beam(d), fixture_lattice(d),
               rstartflg=1)
    (tmp_path / "fort.50").write_bytes(d["ckpt50"].tobytes())
    print(f"io restart fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    worst = _rows_err(tmp_path, d, "b", 9)
    record(test="deck_restart", fused=fused, rows=worst)
    assert worst < TOL


def test_deck_binary_particle_input(tmp_path, gpu):
    from impactz import particle_io as P
    from impactz.beam import write_deck
    from impactz.deck import run_deck
    d = load_golden("io_ckpt")
    grid = tuple(int(v) for v in d["grid"])
    write_deck(str(tmp_path / "ImpactZ.in"), 1000, grid, fixture_beam(d), fixture_lattice(d),
               flagdist=199, distparam=d["c_distparam"])
    P.write_binary_particles(str(tmp_path / "fort.40"), d["pts_in"])
    print("io binary input", flush=True)
    run_deck(str(tmp_path), fused=True)
    worst = _rows_err(tmp_path, d, "c", 17)
    record(test="deck_binary_input", rows=worst)
    assert worst < TOL
