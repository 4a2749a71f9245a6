"""Lattice-scan jobs (izh_job): a bunch stepped through a periodic lattice
with its passes enqueued asynchronously and blocks of whole lattice periods
replayed from a captured HIP graph.  Each job's records and final particles
must equal izh_run's on the same input (1e-10), with and without graphs,
alone and interleaved with other jobs on the same GPU."""
import numpy as np
import pytest

from conftest import col_err, full_row_err
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu


def _setup(gpu, n, grid, gradient, seed, precision=None, period=0):
    from impactz import beam as B
    from impactz.sim import Simulation
    bm = B.Beam()
    pts = B.gaussian_bunch(n, seed=seed)
    lat = B.fodo(1, gradient=gradient)
    sim = Simulation(pts, bm, lat, grid, precision=gpu.FP64 if precision is None else precision)
    sim.bunch.resort_period = period
    return sim


def _period_segments(sim):
    """The izh_segments of one lattice period (the job repeats them)."""
    return sim._segments_of(sim._window(0, 16))


def _solo(gpu, n, grid, gradient, seed, nsteps, precision=None, period=0):
    sim = _setup(gpu, n, grid, gradient, seed, precision, period)
    from impactz.beam import fodo
    sim.lattice = fodo(nsteps // 16 + 1, gradient=gradient)
    rec = sim.run_fused(nsteps)
    rows = [sim.diag(r) for r in rec]
    pts = sim.bunch.getpts()
    z = sim.z
    sim.close()
    return rows, pts, z


def _job(gpu, n, grid, gradient, seed, nsteps, use_graph, precision=None, period=0):
    sim = _setup(gpu, n, grid, gradient, seed, precision, period)
    job = gpu.Job(sim.bunch, sim.field, sim.dom, 0.0, _period_segments(sim), nsteps,
                  sim.beam.current, use_graph)
    return sim, job


@pytest.mark.parametrize("use_graph,period", [(False, 1), (True, 1), (True, 4), (True, 16)])
def test_job_equals_fused_run(gpu, use_graph, period):
    nsteps = 70
    rows_a, pts_a, z_a = _solo(gpu, 30_000, (16, 16, 32), 10.0, 5, nsteps, period=period)
    sim, job = _job(gpu, 30_000, (16, 16, 32), 10.0, 5, nsteps, use_graph, period=period)
    try:
        # budgets of a block and a bit: graph replays at block boundaries,
        # single passes between them
        step = job.info()["block"] + 7 if use_graph else 7
        while job.advance(step) > 0:
            pass
        z, rec = job.finish()
        info = job.info()
        worst = max(full_row_err(sim.diag(rec[k]), rows_a[k]) for k in range(nsteps))
        pe = col_err(sim.bunch.getpts(), pts_a)
        record(test="job_vs_run", use_graph=use_graph, period=period, steps=nsteps,
               moments=worst, particles=pe, **info)
        assert worst < TOL and pe < TOL
        assert z == pytest.approx(z_a, abs=1e-12)
        assert info["graph"] == use_graph
        if use_graph:
            # 70 steps: blocks of whole periods holding an even number of re-sorts
            assert info["block"] == (32 if period == 16 else 16)
            assert info["launches"] >= 1
    finally:
        job.close()
        sim.close()


def test_jobs_interleaved_equal_each_alone(gpu):
    """Three jobs of different sizes, grids, gradients and forms advanced in
    turn a few steps at a time, all in flight on the GPU together: each
    one's results equal its own run alone."""
    specs = [(40_000, (16, 16, 32), 8.0, 11, 48, None),
             (20_000, (32, 32, 32), 12.0, 12, 64, None),
             (60_000, (16, 16, 16), 6.0, 13, 40, "mixed")]
    solos, live = [], []
    for n, grid, gr, seed, steps, form in specs:
        prec = gpu.MIXED if form == "mixed" else gpu.FP64
        solos.append(_solo(gpu, n, grid, gr, seed, steps, prec))
    try:
        for n, grid, gr, seed, steps, form in specs:
            prec = gpu.MIXED if form == "mixed" else gpu.FP64
            live.append(_job(gpu, n, grid, gr, seed, steps, True, prec))
        pending = list(range(len(live)))
        while pending:
            for i in list(pending):
                # 21 passes: a graph block of 16 where one starts, else single passes
                if live[i][1].advance(21) == 0:
                    pending.remove(i)
        for i, ((sim, job), (rows_a, pts_a, _)) in enumerate(zip(live, solos)):
            _, rec = job.finish()
            worst = max(full_row_err(sim.diag(rec[k]), rows_a[k]) for k in range(len(rows_a)))
            pe = col_err(sim.bunch.getpts(), pts_a)
            record(test="jobs_interleaved", job=i, moments=worst, particles=pe,
                   launches=job.info()["launches"])
            assert job.info()["launches"] >= 1
            # the mixed form is deterministic too: the same job alone or in the mix
            assert worst < TOL and pe < TOL, (i, worst, pe)
    finally:
        for sim, job in live:
            job.close()
            sim.close()


def test_job_refuses_rf_and_wakes(gpu):
    from impactz import hip
    sim = _setup(gpu, 5_000, (16, 16, 16), 10.0, 3)
    try:
        segs = _period_segments(sim)
        bad = hip.segment(0.01, 101, [1e6, 324e6, 0, 1, 1], 10, 1.0, 1.0, 0.1, 0.0, None)
        with pytest.raises(hip.IZHError, match="RF cavity"):
            hip.Job(sim.bunch, sim.field, sim.dom, 0.0, [segs[0], bad], 8, sim.beam.current)
        with pytest.raises(hip.IZHError, match="space-charge"):
            hip.Job(sim.bunch, sim.field, sim.dom, 0.0, segs, 8, 0.0)
    finally:
        sim.close()
