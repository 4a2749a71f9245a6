"""The serving layer (impactz.scan.serve): requests of uneven particle counts,
grids and step counts arrive over time, wait for HBM under a budget, run as
jobs side by side and leave as they finish.  Each request's results must be
the same as the same request run alone (1e-10), and the budget must hold."""
import numpy as np
import pytest

from conftest import full_row_err
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu


def test_served_requests_equal_each_alone(gpu):
    from impactz import scan
    from impactz.sim import diag_row
    pool = scan.Pool(400_000, seed=3)
    reqs = [scan.Request(30_000, (16, 16, 16), 24, 8.0, gpu.FP64, 0.00),
            scan.Request(80_000, (16, 16, 32), 40, 12.0, gpu.MIXED, 0.02),
            scan.Request(10_000, (32, 32, 32), 17, 6.0, gpu.FP64, 0.05),
            scan.Request(50_000, (16, 16, 16), 33, 14.0, gpu.MIXED, 0.05),
            scan.Request(20_000, (16, 32, 16), 20, 9.5, gpu.FP64, 0.30),
            scan.Request(60_000, (32, 32, 32), 48, 11.0, gpu.MIXED, 0.31)]
    # room for about two requests at a time: later ones queue for memory
    budget = 2 * max(scan.footprint(r.n, r.grid, r.precision) for r in reqs)
    try:
        out = scan.serve(reqs, pool, budget, keep_records=True,
                         log=lambda m: print(m, flush=True))
        assert sorted(out["records"]) == list(range(len(reqs)))
        assert out["peak_bytes"] <= budget
        assert out["p50_s"] > 0 and out["p99_s"] >= out["p50_s"]
        worst = 0.0
        for i, r in enumerate(reqs):
            # the same request alone (its window of the pool as served)
            lv = scan.Live(r, pool)
            while lv.job.advance(1 << 30) > 0:
                pass
            alone = lv.finish()
            lv.close()
            got = out["records"][i]
            assert got.shape == alone.shape == (r.steps, gpu.RUN_RECORD)
            c = gpu.physconst(pool.beam.freq)
            for k in range(r.steps):
                refpt = np.array([0, 0, 0, 0, alone[k, 45], alone[k, 46]])
                a = diag_row(alone[k, :37], alone[k, 37:43], alone[k, 43], int(alone[k, 44]),
                             refpt, pool.beam.mass, c.scxl, c.pi)
                b = diag_row(got[k, :37], got[k, 37:43], got[k, 43], int(got[k, 44]),
                             refpt, pool.beam.mass, c.scxl, c.pi)
                worst = max(worst, full_row_err(b, a))
        record(test="serve_alone_vs_mix", requests=len(reqs), moments=worst,
               p50=out["p50_s"], p99=out["p99_s"], peak_jobs=out["peak_jobs"])
        assert worst < TOL, worst
    finally:
        pool.close()


def test_footprint_bounds_device_memory(gpu):
    """The admission footprint is at least what a job takes on the device."""
    import ctypes as C
    from impactz import scan
    rt = C.CDLL("libamdhip64.so")
    free0, free1, total = C.c_size_t(), C.c_size_t(), C.c_size_t()
    pool = scan.Pool(1_200_000, seed=4)
    try:
        for n, grid, prec in ((1_000_000, (64, 64, 64), gpu.MIXED),
                              (500_000, (32, 32, 64), gpu.FP64)):
            r = scan.Request(n, grid, 20, 10.0, prec)
            assert rt.hipDeviceSynchronize() == 0
            assert rt.hipMemGetInfo(C.byref(free0), C.byref(total)) == 0
            lv = scan.Live(r, pool)
            while lv.job.advance(1 << 30) > 0:
                pass
            lv.finish()
            assert rt.hipMemGetInfo(C.byref(free1), C.byref(total)) == 0
            used = free0.value - free1.value
            fp = scan.footprint(n, grid, prec)
            record(test="footprint", n=n, grid=grid, used=used, footprint=fp)
            lv.close()
            # the allocator rounds up to its pages: allow 4 MiB per buffer kind
            assert used <= fp + 64 * (4 << 20), (used, fp)
    finally:
        pool.close()
