"""Bends (type 4) on the GPU: the entry points izh_fpol_dipole,
izh_sector_dipole, izh_bpol_dipole and izh_isr_chicane against the oracle,
and the deck path against the reference executable's runs at zero current
(tests/golden/exe_bend.npz, both integrators)."""
import os

import numpy as np
import pytest

from conftest import deck_files_err, fixture_beam, fixture_lattice, load_golden
from test_gpu_parity import record

pytestmark = pytest.mark.gpu

BEND = [0.25, 0.5, 0.0, 0.02, 0.1, 0.15, 0.5, -0.3, 0.4] + [0.0] * 5


def _col_err(got, ref):
    scale = np.abs(ref).max(axis=0) + 1e-300
    return float((np.abs(got - ref) / scale).max())


def test_bend_entry_points_match_oracle(gpu, oracle):
    import ctypes as C
    from impactz import beam as B, hip
    rt = C.CDLL("libamdhip64.so")
    beam = B.Beam(current=0.0)
    pts = B.gaussian_bunch(4000, seed=5, beam=beam)
    ref = pts.copy()
    refpt = np.array([0, 0, 0, 0, 0, -beam.gamma0])
    cst = oracle.physconst(beam.freq)
    b = hip.BeamBunch.from_beam(pts, beam, hip.FP64)
    rd_host = np.random.default_rng(9).standard_normal(pts.shape[0])
    rd = C.c_void_p()
    assert rt.hipMalloc(C.byref(rd), C.c_size_t(rd_host.nbytes)) == 0
    assert rt.hipMemcpy(rd, rd_host.ctypes.data_as(C.c_void_p), C.c_size_t(rd_host.nbytes),
                        1) == 0   # hipMemcpyHostToDevice
    try:
        L, tau, z = 0.8, 0.1, 0.0
        b.fpol_dipole(L, BEND)
        oracle.dipole_face(cst, ref, refpt, L, BEND, False, beam.charge, beam.mass)
        for _ in range(3):
            z = b.sector_dipole(z, tau, L, BEND)
        zo = 0.0
        for _ in range(3):
            zo = oracle.sector_dipole(cst, ref, refpt, zo, tau, L, BEND, beam.charge,
                                      beam.mass)
        # fewer deviates than uploaded particles: refused before any launch
        with pytest.raises(hip.IZHError, match="deviates"):
            b.isr_chicane(tau, L, BEND, rd.value, pts.shape[0] - 1)
        b.isr_chicane(tau, L, BEND, rd.value, pts.shape[0])
        oracle.isr_chicane(cst, ref, refpt, tau, L, BEND, beam.mass, rd_host)
        b.bpol_dipole(L, BEND)
        oracle.dipole_face(cst, ref, refpt, L, BEND, True, beam.charge, beam.mass)
        got = b.getpts()
        err = _col_err(got[:, :6], ref[:, :6])
        record(test="bend_entry_points", case="4000 particles", err=err)
        assert err < 1e-11, err   # FMA contraction on the GPU, none in the oracle
        assert z == zo
        assert b.refpt[4] == pytest.approx(refpt[4], rel=1e-15)
    finally:
        b.close()
        rt.hipFree(rd)


@pytest.mark.parametrize("flagmap", [1, 2])
def test_bend_deck_matches_executable(tmp_path, gpu, flagmap):
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    from impactz import particle_io as P
    d = load_golden("exe_bend")
    grid = tuple(int(v) for v in d["grid"])
    lat = fixture_lattice(d)
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid, fixture_beam(d),
               lat, flagmap=flagmap)
    print(f"bend deck flagmap={flagmap}", flush=True)
    run_deck(str(tmp_path), fused=True)
    worst = 0.0
    for k in (24, 25, 26):
        ref = d[f"m{flagmap}_fort{k}"]
        got = P.read_phase(str(tmp_path / f"fort.{k}"))
        assert got.shape == ref.shape
        scale = np.abs(ref).max(axis=0) + 1e-300
        worst = max(worst, float((np.abs(got - ref) / scale).max()))
    # every column of every diagnostic file, fort.28 and fort.32 exactly
    worst = max(worst, deck_files_err(str(tmp_path), d, f"m{flagmap}_"))
    err = _col_err(P.read_phase(str(tmp_path / "fort.60")), d[f"m{flagmap}_phase60"])
    record(test="bend_deck", case=f"exe_bend flagmap={flagmap}", moments=worst,
           particles=err)
    assert worst < 1e-10 and err < 1e-10, (worst, err)
