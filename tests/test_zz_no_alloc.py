"""izh_diagnostic1_rows, the entry point that allocates nothing: the moment
sums of device rows the caller holds, computed with the library's static
scratch, must equal izh_diagnostic1's on a bunch holding the same rows in the
same order (bitwise: the same kernels sum in the same order), on two calls in
a row, the second with fewer particles, and for an empty input."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rt():
    rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


H2D, D2H = 1, 2


def test_diagnostic1_rows_equals_diagnostic1(gpu):
    from impactz import beam as B
    from impactz import hip
    rt = _rt()
    beam = B.Beam()
    cap = 60_000
    rows_d, out_d = C.c_void_p(), C.c_void_p()
    # the test allocates the caller's buffers; the entry point allocates nothing
    assert rt.hipMalloc(C.byref(rows_d), 6 * cap * 8) == 0
    assert rt.hipMalloc(C.byref(out_d), 44 * 8) == 0
    try:
        for n, seed in ((cap, 21), (17_000, 22), (0, 23)):
            print(f"diagnostic1_rows n={n}", flush=True)
            pts = B.gaussian_bunch(max(n, 1), seed=seed)[:n]
            rows = np.ascontiguousarray(np.zeros((6, cap)))
            rows[:, :n] = pts[:, :6].T
            assert rt.hipMemcpy(rows_d, rows.ctypes.data, rows.nbytes, H2D) == 0
            hip.check(hip.lib().izh_diagnostic1_rows(rows_d, cap, n, out_d, None))
            assert rt.hipDeviceSynchronize() == 0
            got = np.zeros(44)
            assert rt.hipMemcpy(got.ctypes.data, out_d, 44 * 8, D2H) == 0
            if n:
                bunch = hip.BeamBunch.from_beam(pts, beam)
                s, m, r2 = bunch.diagnostic1()
                bunch.close()
                want = np.concatenate([s, m, [r2]])
            else:
                want = np.zeros(44)
            assert np.array_equal(got, want), (n, np.max(np.abs(got - want)))
        # bad sizes are refused before anything is enqueued
        with pytest.raises(hip.IZHError, match="bad argument"):
            hip.check(hip.lib().izh_diagnostic1_rows(rows_d, 10, 20, out_d, None))
        with pytest.raises(hip.IZHError, match="beyond the maximum"):
            hip.check(hip.lib().izh_diagnostic1_rows(rows_d, (1 << 40) + 1, 5, out_d, None))
    finally:
        rt.hipFree(rows_d)
        rt.hipFree(out_d)
