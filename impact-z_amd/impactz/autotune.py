"""The autotuner of the fused run: for a bunch size, grid and form it times
short lattice-scan jobs with each candidate re-sort period (the passes
between two counting-sort scatters of the particles into grid tiles, which
set how far the particles may drift from the tile their gather box and
deposit window were built for) and keeps the fastest.  Results are a table
keyed by form, grid and particle-count decade; `lookup` gives the entry of
the nearest decade for the grid, or the library default.

    python -m impactz.autotune [--out FILE] [--forms fp64,mixed]

measures the table on the GPU (the sizes of the served and batched scans
and of the large single bunch) and writes it as JSON.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

from . import hip

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CANDIDATES = (1, 2, 4, 8)
GRIDS = ((32, 32, 32), (64, 64, 64), (64, 64, 128), (128, 128, 128), (128, 128, 256))
SIZES = (100_000, 1_000_000, 10_000_000, 50_000_000)

_table = None


def _key(precision: int, grid) -> str:
    return ("fp64" if precision == hip.FP64 else "mixed") + ":" + "x".join(map(str, grid))


def measure(pool, n: int, grid, precision: int, period: int, steps: int = 48) -> float:
    """Seconds per step of a job of `steps` steps (graph replays) with the
    given re-sort period, after one untimed job of the same shape."""
    from .scan import Live, Request, hip_sync
    best = math.inf
    for rep in range(2):
        req = Request(n, tuple(grid), steps, 10.0, precision)
        hip_sync()
        t0 = time.perf_counter()
        lv = Live(req, pool, period)
        while lv.job.advance(1 << 30) > 0:
            pass
        lv.finish()
        dt = time.perf_counter() - t0
        lv.close()
        if rep:
            best = min(best, dt / steps)
    return best


def tune(pool, sizes=SIZES, grids=GRIDS, forms=(hip.FP64, hip.MIXED), log=print) -> dict:
    table = {}
    for prec in forms:
        for g in grids:
            for n in sizes:
                if n > pool.n:
                    continue
                times = {p: measure(pool, n, g, prec, p) for p in CANDIDATES}
                best = min(times, key=times.get)
                table.setdefault(_key(prec, g), {})[str(int(round(math.log10(n))))] = {
                    "resort_period": best,
                    "ms_per_step": {str(p): round(1e3 * t, 4) for p, t in times.items()}}
                log(f"{_key(prec, g)} n={n}: " +
                    " ".join(f"K={p} {1e3 * t:.3f} ms" for p, t in times.items()) +
                    f" -> {best}")
    return table


def load(path: str = TABLE) -> dict:
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(path):
            with open(path) as f:
                _table = json.load(f)
    return _table


def lookup(n: int, grid, precision: int) -> int:
    """The re-sort period the table gives the nearest particle-count decade
    on this grid and form; 0 (the library default) without an entry."""
    ent = load().get(_key(precision, tuple(grid)))
    if not ent:
        return 0
    d = math.log10(max(n, 1))
    k = min(ent, key=lambda s: abs(int(s) - d))
    return int(ent[k]["resort_period"])


def step_ms(n: int, grid, precision: int):
    """Measured milliseconds per step at the tuned re-sort period, taken
    from the table's two particle-count decades around n on this grid and
    form (linear in n between them, proportional to n beyond them); None
    without an entry."""
    ent = load().get(_key(precision, tuple(grid)))
    if not ent:
        return None
    # the table's keys are the rounded decades of the measured sizes
    size_of = {str(int(round(math.log10(sz)))): sz for sz in SIZES}
    pts = sorted((float(size_of.get(k, 10.0 ** int(k))), min(v["ms_per_step"].values()))
                 for k, v in ent.items())
    if n <= pts[0][0]:
        return pts[0][1]
    for (n0, t0), (n1, t1) in zip(pts, pts[1:]):
        if n <= n1:
            return t0 + (t1 - t0) * (n - n0) / (n1 - n0)
    n1, t1 = pts[-1]
    return t1 * n / n1


def tuner_for_requests(req) -> int:
    """The tuner hook of impactz.scan: a request's re-sort period."""
    return lookup(req.n, req.grid, req.precision)


def main(argv=None) -> int:
    from .scan import Pool
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--sizes", default=",".join(str(s) for s in SIZES))
    args = ap.parse_args(argv)
    hip.init(0)
    sizes = tuple(int(s) for s in args.sizes.split(","))
    pool = Pool(max(sizes))
    forms = tuple(hip.FP64 if f == "fp64" else hip.MIXED for f in args.forms.split(","))
    table = tune(pool, sizes=sizes, forms=forms, log=lambda m: print(m, flush=True))
    pool.close()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
