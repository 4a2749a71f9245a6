"""izh_run (the fused, sorted, device-resident pipeline) against izh_step and
the reference.  Same bar as the step-by-step path: 1e-10 relative (FP64)."""
import numpy as np
import pytest

from conftest import (col_err, fixture_beam, fixture_lattice, fixture_row_err,
                      full_row_err, load_golden, rel_err)
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu


def _sims(gpu, pts, bm, lat, grid, precision=None):
    from impactz.sim import Simulation
    p = gpu.FP64 if precision is None else precision
    return (Simulation(pts, bm, lat, grid, precision=p),
            Simulation(pts, bm, lat, grid, precision=p))


def test_fused_matches_stepwise_100k_32cube(gpu):
    from impactz import beam as B
    bm = B.Beam()
    pts = B.gaussian_bunch(100_000, seed=1)
    lat = B.fodo(1)
    a, b = _sims(gpu, pts, bm, lat, (32, 32, 32))
    rows_a = []
    for k, e in enumerate(a.segments()):
        rows_a.append(a.diag(a.step(e)))
    rec = b.run_fused(16)
    worst = 0.0
    for k in range(16):
        worst = max(worst, full_row_err(b.diag(rec[k]), rows_a[k]))
    pe = col_err(b.bunch.getpts(), a.bunch.getpts())
    de = rel_err(b.dom.range, a.dom.range)
    record(test="fused_vs_step", case="100k_32cube", steps=16, moments=worst,
           particles=pe, domain=de)
    assert worst < TOL and pe < TOL and de < TOL
    assert abs(b.z - a.z) < 1e-14
    assert b.bunch.nptlocal == a.bunch.nptlocal == 100_000


@pytest.mark.parametrize("name,period", [("fodo_traj", 1), ("fodo_traj", 4),
                                         ("fodo_16cube", 1), ("fodo_aniso", 1),
                                         ("fodo_lossy", 1), ("fodo_lossy", 3),
                                         ("cf_sol", 1), ("cf_sol", 5), ("rf", 1), ("rf", 4),
                                         ("kquad", 1), ("kquad", 4)])
def test_fused_matches_reference(gpu, name, period):
    """izh_run against the reference, re-sorting every pass or every
    period-th pass (the passes between keep the particles in their slots,
    lost ones marked in place)."""
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    sim.bunch.resort_period = period
    assert sim.bunch.resort_period == period
    nsteps = int(d["nsteps"])
    print(f"fused run {name} resort period {period}", flush=True)
    rec = sim.run_fused(nsteps)
    worst = 0.0
    for k in range(nsteps):
        worst = max(worst, fixture_row_err(sim.diag(rec[k]), d, k + 1))
    fin = sim.bunch.getpts()
    assert fin.shape == d["final"].shape
    pe = col_err(fin, d["final"])
    ce = abs(sim.bunch.current / float(d["final_current"]) - 1)
    record(test="fused_vs_reference", case=name, period=period, steps=nsteps,
           moments=worst, particles=pe, current=ce)
    assert worst < TOL and pe < TOL and ce < 1e-14
    # the reference energy of every step (RF cavities change it)
    assert np.allclose(-rec[:, 46], d["fort18"][1:nsteps + 1, 2], rtol=1e-13)
    # Npt of each step's moments follows the reference's losses
    if name == "fodo_lossy":
        assert rec[-1, 44] == d["final"].shape[0]


def test_fused_split_runs_and_mixing_with_steps(gpu):
    """Two runs of 8 == one of 16; a run followed by single steps continues
    the same trajectory."""
    from impactz import beam as B
    bm = B.Beam()
    pts = B.gaussian_bunch(40_000, seed=3)
    lat = B.fodo(1)
    a, b = _sims(gpu, pts, bm, lat, (16, 16, 32))
    ra = a.run_fused(16)
    rb1 = b.run_fused(8)
    rb2 = b.run_fused(8, start=8)
    rb = np.concatenate([rb1, rb2])
    worst = max(full_row_err(b.diag(rb[k]), a.diag(ra[k])) for k in range(16))
    assert worst < 1e-12
    assert col_err(b.bunch.getpts(), a.bunch.getpts()) < 1e-12
    c, _ = _sims(gpu, pts, bm, lat, (16, 16, 32))
    c.run_fused(8)
    segs = list(c.segments())[8:16]
    for e in segs:
        c.step(e)
    assert col_err(c.bunch.getpts(), a.bunch.getpts()) < 1e-12


def test_fused_mixed_precision_error(gpu):
    from impactz import beam as B
    bm = B.Beam()
    pts = B.core_halo_bunch(200_000, seed=1)
    lat = B.fodo(2)
    f64, _ = _sims(gpu, pts, bm, lat, (32, 32, 64))
    mx, _ = _sims(gpu, pts, bm, lat, (32, 32, 64), precision=gpu.MIXED)
    r64 = f64.run_fused(32)
    r32 = mx.run_fused(32)
    rms_cols = [6, 8, 12, 14, 18, 20]
    worst = 0.0
    for k in range(32):
        d64 = f64.diag(r64[k])
        d32 = mx.diag(r32[k])
        worst = max(worst, float(np.max(np.abs(d32[rms_cols] - d64[rms_cols]) /
                                        np.abs(d64[rms_cols]))))
    e64 = np.stack([f64.field.download(w) for w in (gpu.EX, gpu.EY, gpu.EZ)])
    e32 = np.stack([mx.field.download(w) for w in (gpu.EX, gpu.EY, gpu.EZ)])
    fl2 = float(np.linalg.norm(e32 - e64) / np.linalg.norm(e64))
    record(test="fused_mixed_vs_fp64", steps=32, rms=worst, field_l2=fl2)
    assert worst < 1e-4 and fl2 < 1e-3


def test_fused_kmode_quadrupole_with_varying_charge(gpu):
    """The K-mode quadrupole's chromatic map as a push kind of izh_run, on a
    bunch whose particles carry three q/mc^2 (each particle's own strength):
    equal to izh_step, with every re-sort period."""
    from impactz import beam as B
    from impactz.sim import Simulation
    bm = B.Beam()
    pts = B.gaussian_bunch(20_000, seed=2)
    q = bm.charge / bm.mass
    pts[:, 6] = q * np.array([1.0, 0.5, 0.25])[np.arange(pts.shape[0]) % 3]
    lat = [B.Element(0.1, 4, 10, 1, 5.4, -1.0, 1.0), B.Element(0.4, 4, 10, 0, 1.0),
           B.Element(0.1, 4, 10, 1, -5.4, -1.0, 1.0), B.Element(0.4, 4, 10, 0, 1.0)]
    a = Simulation(pts, bm, lat, (16, 16, 16))
    rows_a = [a.diag(a.step(e)) for e in a.segments()]
    pa = a.bunch.getpts()
    for period in (1, 3):
        b = Simulation(pts, bm, lat, (16, 16, 16))
        b.bunch.resort_period = period
        rec = b.run_fused(16)
        worst = max(full_row_err(b.diag(rec[k]), rows_a[k]) for k in range(16))
        pe = col_err(b.bunch.getpts(), pa)
        record(test="fused_kmode_vs_step", period=period, moments=worst, particles=pe)
        assert worst < TOL and pe < TOL
        b.close()
    a.close()
