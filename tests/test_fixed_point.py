"""The deposit's short round-to-integer (izh_field.hip fx_round<true>): for
|x| < 2^50, adding 1.5 * 2^52 in IEEE FP64 (round to nearest even) leaves
round-half-even(x) in the low mantissa bits.  Checked here in numpy's FP64
(the same IEEE add) against np.rint, which rounds half to even, on random
magnitudes up to 2^50, on exact ties and at the bound; the fused deposit's
grid charge is checked bitwise against the exact conversion on the GPU
(test_gpu_configs.py, C3)."""
import numpy as np

MAGIC = np.float64(6755399441055744.0)          # 1.5 * 2^52
MAGIC_BITS = np.int64(0x4338000000000000)


def short_round(x: np.ndarray) -> np.ndarray:
    t = (x.astype(np.float64) + MAGIC).view(np.int64)
    return t - MAGIC_BITS


def test_short_round_equals_round_half_even():
    rng = np.random.default_rng(5)
    mag = np.ldexp(rng.random(2_000_000), rng.integers(0, 51, 2_000_000))
    x = np.where(rng.random(mag.size) < 0.5, -mag, mag)
    ties = np.floor(x[:200_000]) + 0.5
    edges = np.array([0.0, -0.0, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5,
                      0.49999999999999994, 2.0**50 - 0.5, -(2.0**50) + 0.5,
                      2.0**50 - 1.0, 123456789.5, 3.0e14 + 0.25])
    for v in (x, ties, edges):
        assert np.array_equal(short_round(v), np.rint(v).astype(np.int64))


def test_short_round_fails_past_its_bound():
    """The launch only takes the short form below 2^50: at 2^52 the add no
    longer leaves the integer in the low bits, so the guard is needed."""
    v = np.array([2.0**52 + 3.0])
    assert not np.array_equal(short_round(v), np.rint(v).astype(np.int64))
