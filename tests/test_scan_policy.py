"""The served scan's work estimate (impactz.scan.est_seconds over
impactz.autotune.step_ms): the measured step times of the autotuner's
table at the sizes it measured, interpolated between them, and the order it
gives requests."""
import json
import math

import pytest

from impactz import autotune, hip, scan


def _table():
    with open(autotune.TABLE) as f:
        return json.load(f)


def test_step_ms_reads_the_measured_sizes():
    table = _table()
    for key, ent in table.items():
        form, g = key.split(":")
        grid = tuple(int(v) for v in g.split("x"))
        prec = hip.FP64 if form == "fp64" else hip.MIXED
        for size in autotune.SIZES:
            dec = str(int(round(math.log10(size))))
            if dec not in ent:
                continue
            want = min(ent[dec]["ms_per_step"].values())
            assert autotune.step_ms(size, grid, prec) == pytest.approx(want, rel=1e-12)


def test_step_ms_between_and_beyond_the_table():
    grid, prec = (128, 128, 256), hip.FP64
    a, b = autotune.step_ms(1_000_000, grid, prec), autotune.step_ms(10_000_000, grid, prec)
    mid = autotune.step_ms(5_500_000, grid, prec)
    assert min(a, b) <= mid <= max(a, b)
    top = autotune.step_ms(50_000_000, grid, prec)
    assert autotune.step_ms(100_000_000, grid, prec) == pytest.approx(2 * top)
    assert autotune.step_ms(10, (7, 7, 7), prec) is None


def test_est_seconds_orders_requests():
    small = scan.Request(100_000, (32, 32, 32), 50, 0.0, hip.MIXED)
    large = scan.Request(10_000_000, (128, 128, 256), 400, 0.0, hip.MIXED)
    assert scan.est_seconds(small) < scan.est_seconds(large)
    ms = autotune.step_ms(large.n, large.grid, large.precision)
    assert scan.est_seconds(large) == pytest.approx(400 * ms * 1e-3)
    # no table entry: particle-steps at 1e9 per second
    odd = scan.Request(1000, (8, 8, 8), 10, 0.0, hip.FP64)
    assert scan.est_seconds(odd) == pytest.approx(1000 * 10 * 1e-9)


def test_serve_refuses_unknown_policy():
    with pytest.raises(ValueError, match="unknown policy"):
        scan.serve([], None, 0, policy="lifo")
