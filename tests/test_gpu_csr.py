"""The CSR wake (csrwakeTrIGF_FieldQuant, src/Appl/Field.f90:1007-1251) on the
GPU: izh_wakefield of kind 3, one workgroup per mesh line, against the
oracle's restatement (itself pinned to the reference's routine by
tests/golden/csr.npz) on the same particles after conv1st, for bunches
inside a bend, across its exit and past it; and steps with a CSR wake,
granular and fused, against the oracle's steps."""
import numpy as np
import pytest

from conftest import col_err, full_row_err
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu

# an electron beam: 100 MeV, 1 nC at 1.3 GHz
def _beam():
    from impactz.beam import Beam
    return Beam(current=1.3, kinetic=100e6, mass=0.51099895e6, charge=-1.0, freq=1.3e9)


def _bunch(n, seed):
    rng = np.random.default_rng(seed)
    beam = _beam()
    pts = np.zeros((n, 9))
    pts[:, 0] = rng.normal(0, 1e-3, n)
    pts[:, 1] = rng.normal(0, 1e-3, n)
    pts[:, 2] = rng.normal(0, 1e-3, n)
    pts[:, 3] = rng.normal(0, 1e-3, n)
    pts[:, 4] = rng.normal(0, 0.02, n)
    pts[:, 5] = rng.normal(0, 0.05, n)
    pts[:, 6] = beam.charge / beam.mass
    pts[:, 7] = beam.current / beam.freq / n * np.sign(beam.charge)
    pts[:, 8] = np.arange(1, n + 1)
    return beam, pts


@pytest.mark.parametrize("zoff", [0.05, 0.2, 0.3999, 0.45, 1.2])
def test_gpu_csr_wake_matches_oracle(gpu, oracle, zoff):
    hip = gpu
    beam, pts = _bunch(20_000, 7)
    grid = (16, 16, 64)
    b = hip.BeamBunch.from_beam(pts, beam)
    field = hip.FieldQuant(*grid)
    dom = hip.CompDom(*grid)
    pr, _, _ = b.conv1st(0.01, 1.0, 1.0)
    dom.update(pr)
    b.charge(dom, field)
    field.update3O(dom)
    w = hip.wake(3, r0=6.0, zoff=zoff, blength=0.4)
    print(f"csr zoff={zoff}", flush=True)
    b.wakefield(field, dom, w)
    got = field.wake_download()
    tpts = b.getpts()
    info = b.info()
    c = oracle.physconst(beam.freq)
    ow = oracle.make_wake(3, r0=6.0, zoff=zoff, blength=0.4)
    ref = oracle.wakefield(c, tpts, dom.range, grid[2], info.current, info.npt,
                           -info.refpt[5], beam.charge, ow)
    err = float(np.max(np.abs(got[2] - ref[2])) / np.max(np.abs(ref[2])))
    record(test="csr_wake", zoff=zoff, ez=err, scale=float(np.max(np.abs(ref[2]))))
    assert np.all(got[0] == 0) and np.all(got[1] == 0)
    assert np.max(np.abs(ref[2])) > 0
    assert err < 1e-12
    field.close()
    b.close()


def _csr_lattice():
    from impactz.beam import Element
    # drift, a 0.4 m bend of 0.15 rad with v3 = 600 (CSR in the bend and the
    # transient CSR in every element after it), half gap 2 cm, drift, quad
    return [Element(0.1, 2, 1, 0, 1.0),
            Element(0.4, 4, 1, 4, 0.15, 0.0, 600.0, 0.02, more=(0.0, 0.0, 0.0, 0.0, 0.0)),
            Element(0.3, 3, 1, 0, 1.0),
            Element(0.1, 2, 1, 1, 5.0, 0.0, 1.0)]


@pytest.mark.parametrize("fused", [False, True])
def test_gpu_bend_with_current_and_csr_matches_oracle(gpu, oracle, fused):
    """A bend with current in the loop (src/Contrl/AccSimulator.f90:1032-1138,
    1344-1448): space charge with the 100 m bend pipe, the CSR wake inside
    the bend, ISRchicane (the same seeded deviates on both sides), then the
    transient CSR wake in the drift and quadrupole after it (flagbendcsr,
    zwkmin = range(5)/gamma + z - zbleng + sentr).  fused: the steps after
    the bend run through izh_run with a CSR wake per segment."""
    from impactz.sim import Simulation
    beam, pts = _bunch(20_000, 11)
    grid = (16, 16, 32)
    lat = _csr_lattice()
    sim = Simulation(pts, beam, lat, grid)
    run = oracle.Run(pts, beam, lat, grid)
    sim.isr_rng = np.random.default_rng(9)
    run.isr_rng = np.random.default_rng(9)
    steps = list(sim.segments())
    n_before = 2 + 4    # the drift and the bend run step by step
    worst = 0.0
    try:
        rows = []
        for k, e in enumerate(steps[:n_before] if fused else steps):
            print(f"csr lattice step {k + 1} type {e.type}", flush=True)
            rows.append(sim.diag(sim.step(e)))
        if fused:
            recs = sim.run_fused(len(steps) - n_before, start=n_before)
            rows += [sim.diag(r) for r in recs]
        for k, e in enumerate(steps):
            run.step(e)
            worst = max(worst, full_row_err(rows[k], run.diag()))
        pe = col_err(sim.bunch.getpts(), run.particles())
        record(test="csr_lattice", fused=fused, moments=worst, particles=pe)
        assert worst < TOL and pe < TOL, (worst, pe)
        # the CSR wakes ran: the bunch's energy spread changed beyond the
        # zero-current map's (checked in the oracle's own loop)
        assert run.csr["flagbend"] == 1 and run.csr["bendlen"] == 0.4
    finally:
        sim.close()
