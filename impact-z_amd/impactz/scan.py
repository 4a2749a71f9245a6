"""Lattice scans on one GPU: many independent bunches through the FODO
lattice at once, batched (a fixed set started together) or served (requests
arriving over time, admitted by their HBM footprint, continuously batched).

The reference scans a lattice parameter by running its executable once per
point, each a process reading and writing files
(utilities/Engscan.py:28-47, utilities/phaseOptZ.py:212-222).  Here every
scan point is an izh_job: a device-resident bunch, field and domain whose
steps are enqueued on the bunch's own stream, with whole lattice periods
replayed from a captured graph.  The host keeps many jobs in flight: it
enqueues a block of steps of each idle job in turn, polls without waiting,
retires the jobs that finished and starts queued ones as memory frees up.

Synthetic inputs: one resident pool of particles (the 6D Gaussian of the
configurations, truncated at 3 sigma) is uploaded once; a job's bunch is a
window of it copied on the device (izh_bunch_create_from), so starting a
job costs no host sampling or PCIe upload.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import beam as B
from . import hip

STEPS_PER_PERIOD = 16          # FODO: 4 elements x 4 space-charge steps


@dataclasses.dataclass
class Request:
    """One scan point: particles, grid, steps, quadrupole gradient and form;
    arrival time for served requests (s from the start)."""
    n: int
    grid: tuple
    steps: int
    gradient: float
    precision: int = hip.MIXED
    arrival: float = 0.0
    offset: int = 0                # first particle of its window in the pool


def c4_requests(seed: int = 7, nbunch: int = 64, steps: int = 160,
                precision: int = hip.MIXED) -> list[Request]:
    """The batched-launch configuration: particle counts log-uniform in
    [1e5, 4e6], grids from {32^3, 64^3, 64x64x128, 128^3}, each bunch its own
    quadrupole gradient in [5, 15] T/m, from seed 7."""
    rng = np.random.default_rng(seed)
    grids = [(32, 32, 32), (64, 64, 64), (64, 64, 128), (128, 128, 128)]
    out = []
    for _ in range(nbunch):
        n = int(round(math.exp(rng.uniform(math.log(1e5), math.log(4e6)))))
        g = grids[rng.integers(len(grids))]
        out.append(Request(n, g, steps, float(rng.uniform(5.0, 15.0)), precision))
    return out


def c5_requests(seed: int = 11, rate: float = 20.0, duration: float = 300.0,
                precision: int = hip.MIXED) -> list[Request]:
    """The serving configuration: Poisson arrivals at `rate` per second for
    `duration` seconds; particles log-uniform in [1e5, 1e7], a grid from
    32^3 to 128x128x256, 50 to 400 steps and a gradient in [5, 15] T/m."""
    rng = np.random.default_rng(seed)
    grids = [(32, 32, 32), (64, 64, 64), (64, 64, 128), (128, 128, 128), (128, 128, 256)]
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(1e5), math.log(1e7)))))
        g = grids[rng.integers(len(grids))]
        steps = int(rng.integers(50, 401))
        out.append(Request(n, g, steps, float(rng.uniform(5.0, 15.0)), precision, t))
    return out


def footprint(n: int, grid, precision: int) -> int:
    """HBM bytes a job holds: the bunch (both ping-pong copies of the rows,
    index and key words, t positions, the immutable rows, reduction and sort
    scratch) and the field (fixed-point charge, grids, half spectra, Green
    function), as the library allocates them."""
    nx, ny, nz = grid
    pe = 8 if precision == hip.FP64 else 4
    g = nx * ny * nz
    kx = nx + 1
    row = 128 // (2 * pe)           # complex values per 128-byte line
    ldx = (kx + row - 1) // row * row  # x pitch of the half spectra
    tiles = max(1, (nx >> 3) * (ny >> 3) * (nz >> 3))
    bunch = n * (2 * 48 + 4 * 4 + 3 * pe + 3 * 8) + 2048 * 52 * 8 \
        + 2 * (n // 16384 + tiles + 1) * 12 + 3 * tiles * 4
    field = g * (8 + 2 * pe + 4 * pe) + ldx * ny * nz * 2 * pe + ldx * 2 * ny * nz * 2 * pe \
        + (kx + ldx) * (ny + 1) * (nz + 1) * pe + (nx + 2) * (ny + 2) * (nz + 2) * 8
    return int(bunch + field)


class Pool:
    """The resident particle pool jobs take their bunches from."""

    def __init__(self, n: int, seed: int = 1, beam: B.Beam = B.Beam()):
        self.n = n
        self.beam = beam
        self.bunch = hip.BeamBunch.from_beam(B.gaussian_bunch(n, seed=seed, beam=beam), beam)

    def close(self):
        self.bunch.close()


class Live:
    """A job in flight and what it holds."""

    def __init__(self, req: Request, pool: Pool, resort_period: int = 0,
                 use_graph: bool = True):
        self.req = req
        beam = pool.beam
        self.bunch = hip.BeamBunch.from_pool(pool.bunch, req.offset, req.n, beam.current,
                                             req.precision)
        if resort_period:
            self.bunch.resort_period = resort_period
        self.field = hip.FieldQuant(*req.grid, precision=req.precision)
        self.dom = hip.CompDom(*req.grid)
        lat = B.fodo(1, gradient=req.gradient)
        period = [hip.segment(0.5 * e.length / e.nseg, e.type, e.values, e.mapstp,
                              e.radius, e.radius, e.length, 0.0)
                  for e in lat for _ in range(e.nseg)]
        self.job = hip.Job(self.bunch, self.field, self.dom, 0.0, period, req.steps,
                           beam.current, use_graph)
        # a graph replay covers `block` steps: advance by at least that much
        self.block = self.job.info()["block"]
        self.started = time.perf_counter()
        self.result = None

    def finish(self):
        z, rec = self.job.finish()
        self.result = rec
        return rec

    def close(self):
        self.job.close()
        self.field.close()
        self.bunch.close()


def run_batch(reqs: list[Request], pool: Pool, tuner=None, use_graph: bool = True,
              block: int | None = None, keep_records: bool = False) -> dict:
    """Start every request, keep them all in flight, finish them all: the
    batched launch.  Returns timing and the aggregate particle-steps/s over
    the whole batch (job start to the last job's end)."""
    rng = np.random.default_rng(len(reqs))
    for r in reqs:
        r.offset = int(rng.integers(0, pool.n - r.n + 1))
    hip_sync()
    t0 = time.perf_counter()
    live = [Live(r, pool, tuner(r) if tuner else 0, use_graph) for r in reqs]
    blk = block or 2 * STEPS_PER_PERIOD
    pending = list(live)
    while pending:
        for lv in list(pending):
            if lv.job.state() != 0 and lv.job.advance(max(blk, lv.block)) == 0:
                pending.remove(lv)
    recs = [lv.finish() for lv in live]
    dt = time.perf_counter() - t0
    psteps = sum(r.n * r.steps for r in reqs)
    info = [lv.job.info() for lv in live]
    out = {"seconds": dt, "particle_steps": psteps, "value": psteps / dt,
           "jobs": len(reqs), "graph_launches": sum(i["launches"] for i in info)}
    if keep_records:
        out["records"] = recs
    for lv in live:
        lv.close()
    return out


def hip_sync():
    """Wait for the device (between timed phases)."""
    import ctypes as C
    C.CDLL("libamdhip64.so").hipDeviceSynchronize()


def est_seconds(r: Request) -> float:
    """GPU time a request takes alone: its steps at the autotuner table's
    measured step time for its size, grid and form (particle-steps at 1e9/s
    without an entry)."""
    from . import autotune
    ms = autotune.step_ms(r.n, r.grid, r.precision)
    return r.steps * (ms * 1e-3 if ms is not None else r.n * 1e-9)


def serve(reqs: list[Request], pool: Pool, budget_bytes: int, tuner=None,
          use_graph: bool = True, block: int | None = None, keep_records: bool = False,
          log=None, policy: str = "sjf", max_live: int | None = 3,
          aging: float = 0.2) -> dict:
    """Serve requests as they arrive (their `arrival` times, on the wall
    clock from the start): a request waits in the queue until its HBM
    footprint fits the budget (and, with max_live, fewer than max_live jobs
    run), then starts as a job; every idle job gets its next block of steps
    enqueued in turn, so jobs that arrive while others run share the GPU
    with them; finished jobs free their memory at once.  The queue is
    first come first served (policy "fifo") or shortest job first ("sjf":
    the waiting request with the least estimated GPU time, `est_seconds`,
    less `aging` times the seconds it has waited, so long requests are not
    held back for ever).  The defaults (shortest first, ageing 0.2, at most
    3 jobs) were the best of a sweep on the served workload: a large bunch
    fills the device alone, fewer jobs in flight contend less for it, and
    short requests leave first.
    Latency = finish time - arrival time.  Returns p50/p99 latency and the
    particle-steps/s delivered over the session."""
    if policy not in ("fifo", "sjf"):
        raise ValueError(f"serve: unknown policy {policy!r}")
    rng = np.random.default_rng(len(reqs) + 1)
    for r in reqs:
        r.offset = int(rng.integers(0, pool.n - r.n + 1))
    blk = block or 2 * STEPS_PER_PERIOD
    arrivals = sorted(range(len(reqs)), key=lambda i: reqs[i].arrival)
    est = [est_seconds(r) for r in reqs] if policy == "sjf" else None
    queue, live, used = [], [], 0
    lat, done_steps, records = {}, 0, {}
    peak_live, peak_used = 0, 0
    nxt = 0
    hip_sync()
    t0 = time.perf_counter()
    while nxt < len(arrivals) or queue or live:
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and reqs[arrivals[nxt]].arrival <= now:
            queue.append(arrivals[nxt])
            nxt += 1
        # admission by footprint (and the cap on running jobs), in the
        # queue's order
        while queue and (max_live is None or len(live) < max_live):
            qi = 0 if policy == "fifo" else min(
                range(len(queue)),
                key=lambda q: (est[queue[q]] - aging * (now - reqs[queue[q]].arrival),
                               queue[q]))
            ri = queue[qi]
            r = reqs[ri]
            need = footprint(r.n, r.grid, r.precision)
            if used + need > budget_bytes and live:
                break
            queue.pop(qi)
            live.append((ri, Live(r, pool, tuner(r) if tuner else 0, use_graph)))
            used += need
        peak_live = max(peak_live, len(live))
        peak_used = max(peak_used, used)
        progressed = False
        for ri, lv in list(live):
            st = lv.job.state()
            if st == 1:
                lv.job.advance(max(blk, lv.block))
                progressed = True
            elif st == 2:
                rec = lv.finish()
                t_end = time.perf_counter() - t0
                lat[ri] = t_end - reqs[ri].arrival
                done_steps += reqs[ri].n * reqs[ri].steps
                if keep_records:
                    records[ri] = rec
                used -= footprint(reqs[ri].n, reqs[ri].grid, reqs[ri].precision)
                lv.close()
                live.remove((ri, lv))
                progressed = True
                if log:
                    log(f"request {ri} done: {reqs[ri].n} particles, grid {reqs[ri].grid}, "
                        f"{reqs[ri].steps} steps, latency {lat[ri]:.3f} s")
        if not progressed:
            time.sleep(2e-4)
    dt = time.perf_counter() - t0
    v = np.array([lat[i] for i in sorted(lat)])
    out = {"requests": len(reqs), "seconds": dt, "particle_steps": done_steps,
           "value": done_steps / dt,
           "p50_s": float(np.percentile(v, 50)) if len(v) else None,
           "p99_s": float(np.percentile(v, 99)) if len(v) else None,
           "mean_s": float(v.mean()) if len(v) else None,
           "peak_jobs": peak_live, "peak_bytes": peak_used}
    if keep_records:
        out["records"] = records
        out["latency"] = lat
    return out

