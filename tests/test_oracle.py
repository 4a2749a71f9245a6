"""Pins the C oracle (the checker of the HIP path) before it is trusted.

Two kinds of pins:
  * golden vectors written by the compiled reference (tests/golden/*.npz via
    oracle/ref_harness.f90): every stage of a step is fed the reference's own
    input and must return the reference's output;
  * mathematics that fixes a stage independently of the reference (direct
    convolution, Gauss-Legendre cell averages of 1/r, closed-form quadrupole
    and drift maps, CIC partition of unity, frame round trips).
"""
import math

import numpy as np
import pytest

from conftest import (col_err, element_entrance, fixture_beam, fixture_lattice, fort_row,
                      full_row_err, load_golden, rel_err, segment)

CASES = [("fodo_16cube", 1), ("fodo_16cube", 5), ("fodo_aniso", 2),
         ("fodo_lossy", 1), ("fodo_lossy", 2), ("cf_sol", 2), ("cf_sol", 10),
         ("kquad", 3), ("rf", 6), ("rf", 22), ("rf", 30)]


def _rf(oracle, e):
    """The oracle's copy of an RF cavity's field file."""
    return None if e.field is None else oracle.RFData(e.field, fourier=e.type == 105)


@pytest.mark.parametrize("name,step", CASES)
def test_oracle_stages_match_reference(oracle, name, step):
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    c = oracle.physconst(beam.freq)
    tau1 = 0.5 * e.length / e.nseg
    rf, zedge = _rf(oracle, e), element_entrance(d, step)
    # drift1 (first half)
    pts = d[s + "p0"].copy()
    refpt = d[s + "refpt0"].copy()
    z = oracle.map1(c, pts, refpt, float(d[s + "z0"]), tau1, e.type, e.values,
                    e.mapstp, beam.charge, beam.mass, e.length, zedge, rf)
    assert col_err(pts, d[s + "p_map1a"]) < 1e-14
    # conv1st
    pts = d[s + "p_map1a"].copy()
    npt0 = d[s + "p0"].shape[0]
    n, pr, npt, cur = oracle.conv1st(c, pts, refpt, e.radius, e.radius, npt0,
                                     float(d[s + "current0"]))
    assert n == d[s + "p_conv1st"].shape[0] == npt
    assert rel_err(pr, d[s + "ptrange"]) < 1e-15
    assert cur == pytest.approx(float(d[s + "current"]), rel=1e-15)
    assert col_err(pts[:n], d[s + "p_conv1st"]) < 1e-14
    # update_CompDom
    rng, ms = oracle.update_compdom(d[s + "ptrange"], grid)
    assert np.array_equal(rng, d[s + "range"])
    assert np.array_equal(ms, d[s + "msize"])
    # deposit
    rho = oracle.deposit(d[s + "p_conv1st"], grid, d[s + "range"], d[s + "msize"])
    assert rel_err(rho, d[s + "rho"]) < 1e-14
    # open-BC solve (FFT differs from the reference's FFTPACK by rounding only)
    phi = oracle.solve_openbc(d[s + "rho"], d[s + "msize"])
    assert rel_err(phi, d[s + "phi"]) < 1e-13
    # cvbkforth1st
    pts = d[s + "p_conv1st"].copy()
    oracle.cvbkforth1st(c, pts, refpt)
    assert col_err(pts, d[s + "p_cvbk"]) < 1e-14
    # kick1
    pts = d[s + "p_cvbk"].copy()
    oracle.kick1(c, pts, d[s + "phi"], d[s + "range"], d[s + "msize"], refpt,
                 float(d[s + "current"]), 2 * tau1, beam.mass, beam.charge)
    assert col_err(pts, d[s + "p_kick"]) < 1e-13
    # drift1 (second half)
    pts = d[s + "p_kick"].copy()
    z = oracle.map1(c, pts, refpt, z, tau1, e.type, e.values, e.mapstp,
                    beam.charge, beam.mass, e.length, zedge, rf)
    assert col_err(pts, d[s + "p_map1b"]) < 1e-14
    assert z == pytest.approx(float(d[s + "z"]), rel=1e-15)
    assert rel_err(refpt, d[s + "refpt"]) < 1e-15


@pytest.mark.parametrize("name", ["fodo_traj", "fodo_16cube", "fodo_lossy", "cf_sol",
                                  "kquad", "rf"])
def test_oracle_trajectory_matches_reference(oracle, name):
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    nsteps = int(d["nsteps"])
    for k, e in enumerate(run.segments(), 1):
        if k > nsteps:
            break
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            # centroids are compared on the scale of the rms size
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
        assert row[0] == pytest.approx(d["fort18"][k, 1], rel=1e-14)
        # every column of fort.18, 27, 29 and 30 as well, each on its own scale
        assert full_row_err(row, fort_row(d, k)) < 1e-11, k
    fin = d["final"]
    assert run.n == fin.shape[0]
    assert col_err(run.particles(), fin) < 1e-12
    assert run.st.current == pytest.approx(float(d["final_current"]), rel=1e-14)


# ------------------------------------------------------------------ maths

def test_solve_fft_equals_direct_sum(oracle):
    rng = np.random.default_rng(3)
    for grid, h in (((4, 8, 8), (1.0, 0.7, 2.3)), ((8, 4, 16), (0.3, 0.5, 0.9))):
        nx, ny, nz = grid
        rho = rng.standard_normal((nz, ny, nx))
        a = oracle.solve_openbc(rho, h)
        b = oracle.solve_direct(rho, h)
        assert rel_err(a, b) < 1e-13


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def test_igf_is_cell_average_of_inverse_r(oracle):
    """g(i,j,k) = (1/V) * integral of 1/r over the cell centred at (ih, jh, kh)."""
    h = (0.7, 1.1, 1.9)
    g = oracle.green_igf((4, 4, 4), h)
    for (i, j, k) in ((1, 0, 0), (3, 2, 1), (0, 4, 2), (2, 2, 2)):
        xs, wx = _gl(24, (i - 0.5) * h[0], (i + 0.5) * h[0])
        ys, wy = _gl(24, (j - 0.5) * h[1], (j + 0.5) * h[1])
        zs, wz = _gl(24, (k - 0.5) * h[2], (k + 0.5) * h[2])
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        W = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
        ref = np.sum(W / np.sqrt(X ** 2 + Y ** 2 + Z ** 2)) / (h[0] * h[1] * h[2])
        assert g[k, j, i] == pytest.approx(ref, rel=1e-9)


def test_igf_self_cell_unit_cube(oracle):
    """Integral of 1/r over the unit cube about its centre (about 2.38008).

    Written as 6 faces * (1/4) * integral over a face of 1/r (divergence of
    r_hat/r over the cube), a smooth integrand for Gauss-Legendre.
    """
    ys, wy = _gl(64, -0.5, 0.5)
    Y, Z = np.meshgrid(ys, ys, indexing="ij")
    face = np.sum(wy[:, None] * wy[None, :] / np.sqrt(0.25 + Y ** 2 + Z ** 2))
    exact = 6 * 0.25 * face
    assert exact == pytest.approx(2.38008, abs=1e-5)
    g = oracle.green_igf((4, 4, 4), (1.0, 1.0, 1.0))
    assert g[0, 0, 0] == pytest.approx(exact, rel=1e-12)


def test_quad_map_closed_form(oracle):
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    tau, grad = 0.05, 10.0
    refpt = np.zeros(6)
    refpt[5] = -b.gamma0
    xm, r2 = oracle.quad_map(c, tau, grad, 10, refpt, b.charge, b.mass)
    gb = b.gambet0
    brho = gb * b.mass / (299792458.0 * b.charge)
    k = math.sqrt(grad / brho)
    L = tau
    sx = c.scxl
    # (x/Scxl, gamma beta x') transport of a focusing (x) and defocusing (y) plane
    exp_x = np.array([[math.cos(k * L), math.sin(k * L) / (k * gb * sx)],
                      [-k * gb * sx * math.sin(k * L), math.cos(k * L)]])
    exp_y = np.array([[math.cosh(k * L), math.sinh(k * L) / (k * gb * sx)],
                      [k * gb * sx * math.sinh(k * L), math.cosh(k * L)]])
    assert np.allclose(xm[0:2, 0:2], exp_x, rtol=1e-11, atol=0)
    assert np.allclose(xm[2:4, 2:4], exp_y, rtol=1e-11, atol=0)
    assert xm[4, 5] == pytest.approx(L / (sx * gb ** 3), rel=1e-12)
    assert xm[4, 4] == pytest.approx(1.0, abs=1e-15)
    for blk in range(3):
        m = xm[2 * blk:2 * blk + 2, 2 * blk:2 * blk + 2]
        assert np.linalg.det(m) == pytest.approx(1.0, abs=1e-13)
    assert r2[4] == pytest.approx(L / (sx * b.beta0), rel=1e-13)


def _blk(k, L, gb, sx, p=1):
    """Transport of (q/Scxl, (gamma beta)^p q') through a constant spring k^2."""
    if k > 0:
        return np.array([[math.cos(k * L), math.sin(k * L) / (k * gb ** p * sx)],
                         [-k * gb ** p * sx * math.sin(k * L), math.cos(k * L)]])
    return np.array([[1.0, L / (gb ** p * sx)], [0.0, 1.0]])


def test_constfoc_map_closed_form(oracle):
    """Constant focusing (src/Appl/ConstFoc.f90): three harmonic blocks."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    tau, k2 = 0.05, (30.0, 20.0, 5.0)
    refpt = np.zeros(6)
    refpt[5] = -b.gamma0
    xm, r2 = oracle.element_map(c, tau, 2, [*k2, 1.0], 10, refpt, b.charge, b.mass)
    gb, sx = b.gambet0, c.scxl
    assert np.allclose(xm[0:2, 0:2], _blk(math.sqrt(k2[0]), tau, gb, sx), rtol=1e-11, atol=0)
    assert np.allclose(xm[2:4, 2:4], _blk(math.sqrt(k2[1]), tau, gb, sx), rtol=1e-11, atol=0)
    # the longitudinal block scales with (gamma beta)^3
    assert np.allclose(xm[4:6, 4:6], _blk(math.sqrt(k2[2]), tau, gb, sx, 3), rtol=1e-11,
                       atol=0)
    assert np.count_nonzero(xm[0:2, 2:6]) == 0 and np.count_nonzero(xm[2:4, 4:6]) == 0
    assert r2[4] == pytest.approx(tau / (sx * b.beta0), rel=1e-13)


def test_solenoid_map_closed_form(oracle):
    """The solenoid map of src/Appl/Sol.f90: both transverse planes focused
    by (B/2Brho)^2, times the rotation by -B/(2 Brho) * tau."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    tau, b0 = 0.05, 1.5
    refpt = np.zeros(6)
    refpt[5] = -b.gamma0
    xm, _ = oracle.element_map(c, tau, 3, [b0, 0.0, 1.0, 0.0], 10, refpt, b.charge,
                               b.mass)
    gb, sx = b.gambet0, c.scxl
    brho = gb * b.mass / (299792458.0 * b.charge)
    f = _blk(0.5 * b0 / brho, tau, gb, sx)
    a = -b0 / brho / 2 * tau
    rot = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    assert np.allclose(xm[0:4, 0:4], np.kron(rot, f), rtol=1e-11, atol=1e-15)
    assert np.allclose(xm[4:6, 4:6], _blk(0.0, tau, gb, sx, 3), rtol=1e-12, atol=0)


def test_kmode_quad_equals_gradient_map_on_energy(oracle):
    """transfmapK_Quadrupole and the RK6 map of maplinear_Quadrupole are two
    routines of the reference; for on-energy particles (pt = 0) with K = G/Brho
    they must give the same transverse transport."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    rng = np.random.default_rng(4)
    pts = np.zeros((200, 9))
    pts[:, 0] = rng.normal(0, 1e-3, 200) / c.scxl
    pts[:, 2] = rng.normal(0, 1e-3, 200) / c.scxl
    pts[:, 6] = b.charge / b.mass
    brho = b.gambet0 * b.mass / (299792458.0 * b.charge)
    for kq in (5.4, -5.4):
        p1, p2 = pts.copy(), pts.copy()
        r1 = np.zeros(6); r1[5] = -b.gamma0
        r2 = r1.copy()
        oracle.map1(c, p1, r1, 0.0, 0.05, 1, [kq, -1.0, 1.0, 0.0], 10, b.charge, b.mass)
        oracle.map1(c, p2, r2, 0.0, 0.05, 1, [kq * brho, 0.0, 1.0, 0.0], 10, b.charge,
                    b.mass)
        assert col_err(p1[:, :4], p2[:, :4]) < 1e-11
        assert np.all(np.abs(p1[:, 4]) < 1e-15)  # 1/beta_z - 1/beta0 rounds to 1e-16
        assert r1[4] == pytest.approx(r2[4], rel=1e-13)


def test_unsupported_branches_are_refused(oracle):
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    p = np.zeros((1, 9))
    r = np.zeros(6); r[5] = -b.gamma0
    for bt, v in ((1, [1.0, -20.0, 1.0, 0.0]), (1, [1.0, 1.0, 1.0, 0.0]),
                  (2, [1.0, 1.0, 1.0, -1.0]), (4, [0.0] * 4)):
        with pytest.raises(ValueError):
            oracle.map1(c, p, r, 0.0, 0.05, bt, v, 10, b.charge, b.mass)


def test_drift_time_reversal(oracle):
    from impactz.beam import Beam, gaussian_bunch
    b = Beam()
    c = oracle.physconst(b.freq)
    pts = gaussian_bunch(500, seed=5)
    p = pts.copy()
    refpt = np.zeros(6)
    refpt[5] = -b.gamma0
    z = oracle.map1(c, p, refpt, 0.0, 0.2, 0, 0.0, 1, b.charge, b.mass)
    z = oracle.map1(c, p, refpt, z, -0.2, 0, 0.0, 1, b.charge, b.mass)
    assert col_err(p, pts) < 1e-13
    assert abs(z) < 1e-16 and abs(refpt[4]) < 1e-12


def test_deposit_partition_of_unity(oracle):
    rng = np.random.default_rng(1)
    n = 2000
    pts = np.zeros((n, 9))
    pts[:, 0] = rng.uniform(-1, 1, n)
    pts[:, 2] = rng.uniform(-2, 2, n)
    pts[:, 4] = rng.uniform(-3, 3, n)
    pts[:, 7] = rng.uniform(0.5, 1.5, n)
    pr = np.array([-1, 1, -2, 2, -3, 3.0])
    grid = (8, 16, 8)
    rng_, ms = oracle.update_compdom(pr, grid)
    rho = oracle.deposit(pts, grid, rng_, ms)
    assert rho.sum() == pytest.approx(pts[:, 7].sum(), rel=1e-13)
    # a single particle spreads exactly its trilinear weights
    one = np.zeros((1, 9))
    one[0, [0, 2, 4, 7]] = [0.1, -0.3, 0.7, 2.0]
    rho1 = oracle.deposit(one, grid, rng_, ms)
    fx = (0.1 - rng_[0]) / ms[0]
    fy = (-0.3 - rng_[2]) / ms[1]
    fz = (0.7 - rng_[4]) / ms[2]
    i, j, k = int(fx), int(fy), int(fz)
    ax, ay, az = fx - i, fy - j, fz - k
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = ((ax if di else 1 - ax) * (ay if dj else 1 - ay) *
                     (az if dk else 1 - az) * 2.0)
                assert rho1[k + dk, j + dj, i + di] == pytest.approx(w, rel=1e-12, abs=1e-15)
    assert np.count_nonzero(rho1) == 8


def test_frame_round_trip_with_zero_field(oracle):
    """conv1st -> cvbkforth1st -> kick1 with phi = 0 restores the z-frame
    particle (phase wrapped), the identity the reference's frame changes obey."""
    from impactz.beam import Beam, gaussian_bunch
    b = Beam()
    c = oracle.physconst(b.freq)
    pts = gaussian_bunch(300, seed=9)
    pts[:5, 4] += 2 * np.pi * np.array([1, -1, 2, 3, -2])
    p = pts.copy()
    refpt = np.zeros(6)
    refpt[5] = -b.gamma0
    n, pr, _, _ = oracle.conv1st(c, p, refpt, 1.0, 1.0, 300, 0.1)
    grid = (8, 8, 8)
    rng_, ms = oracle.update_compdom(pr, grid)
    oracle.cvbkforth1st(c, p, refpt)
    oracle.kick1(c, p, np.zeros((8, 8, 8)), rng_, ms, refpt, 0.1, 0.01, b.mass, b.charge)
    want = pts.copy()
    pilc = float(np.float32(np.pi))
    nt = np.trunc(want[:, 4] / pilc)
    want[:, 4] = want[:, 4] - nt * pilc - np.fmod(nt, 2) * pilc
    assert col_err(p, want) < 1e-12


# ------------------------------------------ nonlinear integrator (Flagmap=2)

NL_CASES = [("nl_fodo", 1), ("nl_fodo", 5), ("nl_cf_sol", 2), ("nl_cf_sol", 10),
            ("nl_cf_sol", 11), ("nl_rf", 6), ("nl_rf", 22), ("nl_solrf", 7)]


def _zedge(d, step):
    """Entrance z of the element holding the 1-based step."""
    from conftest import fixture_lattice
    z, k = 0.0, 0
    for e in fixture_lattice(d):
        if k + e.nseg >= step:
            return z
        k += e.nseg
        z += e.length
    raise ValueError(step)


@pytest.mark.parametrize("name,step", NL_CASES)
def test_oracle_nonlinear_stages_match_reference(oracle, name, step):
    d = load_golden(name)
    assert int(d["flagmap"]) == 2
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    c = oracle.physconst(beam.freq)
    tau1 = 0.5 * e.length / e.nseg
    pts = d[s + "p0"].copy()
    refpt = d[s + "refpt0"].copy()
    z = oracle.drift2(c, pts, refpt, float(d[s + "z0"]), 2 * tau1)
    assert col_err(pts, d[s + "p_map1a"]) < 1e-14
    pts = d[s + "p_cvbk"].copy()
    refpt_k = refpt.copy()
    rf = None if e.field is None else oracle.RFData(e.field, fourier=True)
    oracle.kick2(c, pts, d[s + "phi"], d[s + "range"], d[s + "msize"], refpt_k,
                 float(d[s + "current"]), 2 * tau1, beam.mass, beam.charge, e.type,
                 e.values, e.length, _zedge(d, step), z, rf)
    assert col_err(pts, d[s + "p_kick"]) < 1e-13
    pts = d[s + "p_kick"].copy()
    z = oracle.drift2(c, pts, refpt_k, z, 2 * tau1)
    assert col_err(pts, d[s + "p_map1b"]) < 1e-14
    assert z == pytest.approx(float(d[s + "z"]), rel=1e-15)
    assert rel_err(refpt_k, d[s + "refpt"]) < 1e-15


@pytest.mark.parametrize("name", ["nl_fodo", "nl_cf_sol", "nl_rf", "nl_solrf"])
def test_oracle_nonlinear_trajectory_matches_reference(oracle, name):
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    nsteps = int(d["nsteps"])
    for k, e in enumerate(run.segments(), 1):
        if k > nsteps:
            break
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
    assert col_err(run.particles(), d["final"]) < 1e-12


@pytest.mark.parametrize("tag", ["sw", "nl"])
def test_oracle_matches_executable_lorentz_runs(oracle, tag):
    """The reference executable's own runs (tests/golden/exe_nonlinear.npz):
    Flagmap=1 switched to 2 by a -25 element, and a Flagmap=2 FODO period.
    Pins the integrator switch and the zedge bookkeeping of the run loop."""
    from impactz.beam import Element
    d = load_golden("exe_nonlinear")
    grid = tuple(int(v) for v in d["grid"])
    lat = [Element(float(r[0]), int(r[1]), int(r[2]), int(r[3]),
                   *(float(v) for v in r[4:8])) for r in d[f"{tag}_lattice"]]
    run = oracle.Run(d["pts_in"], fixture_beam(d), [e for e in lat if e.type >= 0], grid,
                     flagmap=int(d[f"{tag}_flagmap"]))
    k = 0
    for e in lat:
        if e.type == -25:
            run.flagmap = e.mapstp
            continue
        for _ in range(e.nseg):
            run.step(e)
            k += 1
            row = run.diag()
            for f, lo in ((24, 5), (25, 11), (26, 17)):
                ref = d[f"{tag}_fort{f}"][k, 1:]
                got = row[lo:lo + 6]
                scale = np.abs(ref).copy()
                scale[0] = max(scale[0], abs(ref[1]))
                scale[2] = max(scale[2], abs(ref[3]))
                assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
            assert full_row_err(row, fort_row(d, k, f"{tag}_")) < 1e-11, k
            assert run.st.z == pytest.approx(d[f"{tag}_fort24"][k, 0], rel=1e-14)
    assert k == d[f"{tag}_fort24"].shape[0] - 1


def test_drift2_equals_exact_drift(oracle):
    """drift2(2 tau) and the exact drift1(tau) are the same map."""
    from impactz.beam import Beam, gaussian_bunch
    b = Beam()
    c = oracle.physconst(b.freq)
    pts = gaussian_bunch(500, seed=8)
    p1, p2 = pts.copy(), pts.copy()
    r1 = np.zeros(6); r1[5] = -b.gamma0
    r2 = r1.copy()
    z1 = oracle.map1(c, p1, r1, 0.0, 0.05, 0, 0.0, 1, b.charge, b.mass)
    z2 = oracle.drift2(c, p2, r2, 0.0, 0.1)
    assert col_err(p1, p2) < 1e-13 and z1 == z2
    assert r1[4] == pytest.approx(r2[4], rel=1e-14)


def test_kick2_lorentz_push_properties(oracle):
    """Without space charge: no field leaves the particles alone; a pure
    magnetic field (quadrupole, solenoid) keeps |p|; the quadrupole's
    transverse kick matches the linear map for a short step."""
    from impactz.beam import Beam, gaussian_bunch
    b = Beam()
    c = oracle.physconst(b.freq)
    pts = gaussian_bunch(400, seed=9)
    refpt = np.zeros(6); refpt[5] = -b.gamma0
    phi = np.zeros((4, 4, 4))
    rng, ms = np.zeros(6), np.ones(3)
    p = pts.copy()
    oracle.kick2(c, p, phi, rng, ms, refpt.copy(), 0.0, 0.01, b.mass, b.charge, 0,
                 [1.0], 0.1, 0.0, 0.05)
    assert np.array_equal(p, pts)
    gam0 = b.gamma0

    def pmag(q):
        return np.sqrt((gam0 - q[:, 5]) ** 2 - 1.0)  # |gamma beta| from the energy

    for bt, v in ((1, [10.0, 0.0, 1.0, 0.0]), (3, [1.5, 0.0, 0.01, 0.0])):
        p = pts.copy()
        oracle.kick2(c, p, phi, rng, ms, refpt.copy(), 0.0, 0.01, b.mass, b.charge, bt,
                     v, 0.1, 0.0, 0.05)
        # the energy coordinate is unchanged by a magnetic field
        assert np.max(np.abs(p[:, 5] - pts[:, 5])) < 1e-15
        assert np.allclose(pmag(p), pmag(pts), rtol=1e-15)
    # quadrupole: dpx = -G/Brho * x * (gamma beta) * tau for a short step
    p = pts.copy()
    g, tau = 10.0, 1e-4
    oracle.kick2(c, p, phi, rng, ms, refpt.copy(), 0.0, tau, b.mass, b.charge, 1,
                 [g, 0.0, 1.0, 0.0], 0.1, 0.0, 0.05)
    qm = b.charge / b.mass
    brho = b.gambet0 * b.mass / (299792458.0 * b.charge)
    dpx = (p[:, 1] - pts[:, 1])
    want = -g * (pts[:, 0] * c.scxl) * tau / brho * b.gambet0
    assert np.allclose(dpx, want, rtol=2e-2, atol=1e-12)
    del qm


def _const_table(e0=1.0):
    return np.array([[0.0, e0, 0.0, 0.0], [1.0, e0, 0.0, 0.0]])


@pytest.mark.parametrize("bitype", [101, 102, 103, 104])
def test_rf_map_without_field_is_the_linear_drift(oracle, bitype):
    """Field scale 0: the RF cavity's map is the field-free linear map (the
    constant-focusing map with k = 0) and the energy stays."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    refpt = np.array([0, 0, 0, 0, 0.7, -b.gamma0])
    rf = oracle.RFData(_const_table(), fourier=False)
    v = [0.0, 324e6, 25.0, 1.0, 1.0]
    xm, r1 = oracle.element_map(c, 0.05, bitype, v, 10, refpt, b.charge, b.mass, 0.3, 0.5,
                                0.2, rf)
    xd, r2 = oracle.element_map(c, 0.05, 2, [0.0, 0.0, 0.0, 1.0], 10, refpt, b.charge,
                                b.mass)
    assert np.allclose(xm, xd, rtol=1e-15, atol=0)
    assert r1[5] == refpt[5] and r1[4] == pytest.approx(r2[4], rel=1e-15)


@pytest.mark.parametrize("bitype", [101, 103])
def test_rf_energy_gain_in_a_uniform_field(oracle, bitype):
    """A uniform on-axis field E0 at a frequency low enough that the phase
    stays theta0: gamma grows by q E0 cos(theta0) L / mc^2 (theta0 with the
    reference's single-precision asin(1.0), a 1e-8 effect), and every 2x2
    block of the map keeps determinant 1 (the Wronskian of q' = p/Scxl,
    p' = -s q; the entrance and exit terms are shears)."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    e0, th, L = 3.0e6, 30.0, 0.4
    refpt = np.array([0, 0, 0, 0, 0.7, -b.gamma0])
    rf = oracle.RFData(_const_table(), fourier=False)
    v = [e0, 1.0e-3, th, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0]
    xm, r = oracle.element_map(c, L, bitype, v, 40, refpt, b.charge, b.mass, 0.0, L, 0.0, rf)
    # theta0 = Param(4)*asin(1.0)/90 with the REAL(4) asin of the reference
    half_pi = float(np.arcsin(np.float32(1.0)))
    dgam = (b.charge / b.mass) * e0 * math.cos(th * half_pi / 90) * L
    assert -r[5] - b.gamma0 == pytest.approx(dgam, rel=1e-12)
    for blk in range(3):
        m = xm[2 * blk:2 * blk + 2, 2 * blk:2 * blk + 2]
        assert np.linalg.det(m) == pytest.approx(1.0, abs=1e-12)
    assert np.all(xm[:2, 2:] == 0) and np.all(xm[2:4, :2] == 0)


# ------------------------------------------------------------- wakefields

WAKE_CASES = [("wake_file", 2, 2, 1.5), ("wake_file", 12, 2, 0.8), ("wake_solrf", 6, 1, 1.0)]


def _wake_of(oracle, d, step, kind, scale):
    lat = fixture_lattice(d)
    if kind == 2:
        sw = [e for e in lat if e.type == -41]
        btw = 2 if step > 8 else 0
        return oracle.make_wake(2, table=sw[0].field, flagbtw=btw, scale=scale)
    e = segment(d, step)
    v = e.values
    return oracle.make_wake(1, v[11], v[12], v[13], scale=scale)


@pytest.mark.parametrize("name,step,kind,scale", WAKE_CASES)
def test_oracle_wake_stages_match_reference(oracle, name, step, kind, scale):
    """The wake fields from the particles after conv1st (with gamma after
    the first half map, as the run loop has it) and kick1wake from the
    reference's own inputs."""
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    c = oracle.physconst(beam.freq)
    tau1 = 0.5 * e.length / e.nseg
    pts = d[s + "p0"].copy()
    refpt = d[s + "refpt0"].copy()
    oracle.map1(c, pts, refpt, float(d[s + "z0"]), tau1, e.type, e.values, e.mapstp,
                beam.charge, beam.mass, e.length, element_entrance(d, step), _rf(oracle, e))
    w = _wake_of(oracle, d, step, kind, scale)
    got = oracle.wakefield(c, d[s + "p_conv1st"], d[s + "range"], grid[2],
                           float(d[s + "current"]), d[s + "p_conv1st"].shape[0],
                           -refpt[5], beam.charge, w)
    ref = d[s + "wake"]
    assert np.abs(ref).max() > 0
    for k in range(3):
        assert np.abs(got[k] - ref[k]).max() <= 1e-13 * max(np.abs(ref[k]).max(), 1e-300)
    pk = d[s + "p_cvbk"].copy()
    oracle.kick1wake(c, pk, d[s + "phi"], d[s + "range"], d[s + "msize"], refpt,
                     float(d[s + "current"]), 2 * tau1, beam.mass, beam.charge, ref)
    assert col_err(pk, d[s + "p_kick"]) < 1e-13
    # without the wake the kick differs (the wake is not negligible here)
    pn = d[s + "p_cvbk"].copy()
    oracle.kick1(c, pn, d[s + "phi"], d[s + "range"], d[s + "msize"], refpt,
                 float(d[s + "current"]), 2 * tau1, beam.mass, beam.charge)
    assert col_err(pn, d[s + "p_kick"]) > 1e-9


@pytest.mark.parametrize("name", ["wake_file", "wake_solrf"])
def test_oracle_wake_trajectory_matches_reference(oracle, name):
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    nsteps = int(d["nsteps"])
    for k, e in enumerate(run.segments(), 1):
        if k > nsteps:
            break
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
    assert col_err(run.particles(), d["final"]) < 1e-12


def test_wake_particle_on_the_last_mesh_line(oracle):
    """A particle at z = range(6), on the last mesh line: the reference's loop
    gives it weight 1 on line Nz (and its zero upper weight one row past the
    array), so the line density of particles on lines 1..Nz-1 (0-based) is
    uniform up to and including the last line."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    nz = 16
    tab = np.array([[0.0, 5.0e13, 1.0e18, 2.0e18], [1.0, 5.0e13, 1.0e18, 2.0e18]])
    w = oracle.make_wake(2, table=tab)
    rng = np.array([-1, 1, -1, 1, 0.0, 0.15])
    hz = (rng[5] - rng[4]) / (nz - 1)
    pts = np.zeros((15, 9))
    pts[:, 4] = rng[4] + hz * np.arange(1, 16)
    pts[-1, 4] = rng[5]
    out = oracle.wakefield(c, pts, rng, nz, b.current, 15, 2.0, b.charge, w)
    rho = b.current / b.freq / 15 / (hz / 2.0)
    lines = np.zeros(nz)
    lines[1:16] = rho
    want = np.array([-(0.5 * lines[k] + lines[k + 1:].sum()) * 5.0e13 * hz / 2.0
                     for k in range(nz)])
    assert np.abs(out[2] - want).max() <= 1e-13 * np.abs(want).max()


def test_wake_convolution_closed_forms(oracle):
    """A constant longitudinal wake W0 over a bunch of uniform line density:
    E_z(k) = -W0 hz (rho/2 + rho * (number of bins ahead)); the transverse
    wake of a bunch on axis is zero, and it is linear in the offset."""
    from impactz.beam import Beam
    b = Beam()
    c = oracle.physconst(b.freq)
    nz = 16
    tab = np.array([[0.0, 5.0e13, 1.0e18, 2.0e18], [1.0, 5.0e13, 1.0e18, 2.0e18]])
    w = oracle.make_wake(2, table=tab)
    # particles exactly on the mesh lines 1..14 (t frame z), one per line
    rng = np.array([-1, 1, -1, 1, 0.0, 0.15])
    hz = (rng[5] - rng[4]) / (nz - 1)
    pts = np.zeros((14, 9))
    pts[:, 4] = rng[4] + hz * np.arange(1, 15)
    out = oracle.wakefield(c, pts, rng, nz, b.current, 14, 2.0, b.charge, w)
    rho = b.current / b.freq / 14 / (hz / 2.0)
    lines = np.zeros(nz)
    lines[1:15] = rho
    want = np.array([-(0.5 * lines[k] + lines[k + 1:].sum()) * 5.0e13 * hz / 2.0
                     for k in range(nz)])
    # particles on mesh lines may spill a rounding-level weight to a neighbour
    assert np.abs(out[2] - want).max() <= 1e-13 * np.abs(want).max()
    assert np.all(out[0] == 0) and np.all(out[1] == 0)
    pts[:, 0] = 1e-3
    out2 = oracle.wakefield(c, pts, rng, nz, b.current, 14, 2.0, b.charge, w)
    pts[:, 0] = 2e-3
    out3 = oracle.wakefield(c, pts, rng, nz, b.current, 14, 2.0, b.charge, w)
    assert np.abs(out3[0] - 2 * out2[0]).max() <= 1e-13 * np.abs(out3[0]).max()
    assert np.abs(out2[0]).max() > 0


def test_oracle_matches_executable_wake_run(oracle):
    """The reference executable's own run of the -41 wake lattice
    (tests/golden/exe_wake.npz): the wake switches of its element loop and
    the read-in wake end to end."""
    d = load_golden("exe_wake")
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    k = 0
    for k, e in enumerate(run.segments(), 1):
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
        assert full_row_err(row, fort_row(d, k)) < 1e-11, k
    assert k == d["fort24"].shape[0] - 1


def test_oracle_matches_reference_example1(oracle):
    """The reference's examples/Example1 as its executable ran it
    (tests/golden/example1.npz): H- through DTL cells on their Fourier fields
    with the Lorentz integrator; the first 130 of its 1002 steps (the first
    cell and the start of the second) to keep the CPU suite short."""
    d = load_golden("example1")
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    for k, e in enumerate(run.segments(), 1):
        if k > 130:
            break
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-10 * scale), (k, f, got, ref)
        assert run.st.refpt[5] == pytest.approx(-d["fort18"][k, 2], rel=1e-13)
        assert full_row_err(row, fort_row(d, k)) < 1e-10, k


@pytest.mark.parametrize("name,step,kind", [("nl_wake_file", 6, 2), ("nl_wake_solrf", 7, 1)])
def test_oracle_kick2wake_matches_reference(oracle, name, step, kind):
    """Flagmap=2 with a wake: the wake fields after drift2 and kick2wake from
    the reference's own inputs (tests/golden/nl_wake_*.npz)."""
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    c = oracle.physconst(beam.freq)
    tau1 = 0.5 * e.length / e.nseg
    pts = d[s + "p0"].copy()
    refpt = d[s + "refpt0"].copy()
    z = oracle.drift2(c, pts, refpt, float(d[s + "z0"]), 2 * tau1)
    w = _wake_of(oracle, d, step, kind, 1.5 if kind == 2 else 1.0)
    got = oracle.wakefield(c, d[s + "p_conv1st"], d[s + "range"], grid[2],
                           float(d[s + "current"]), d[s + "p_conv1st"].shape[0],
                           -refpt[5], beam.charge, w)
    ref = d[s + "wake"]
    for k in range(3):
        assert np.abs(got[k] - ref[k]).max() <= 1e-13 * max(np.abs(ref[k]).max(), 1e-300)
    pk = d[s + "p_cvbk"].copy()
    rf = None if e.field is None else oracle.RFData(e.field, fourier=True)
    oracle.kick2(c, pk, d[s + "phi"], d[s + "range"], d[s + "msize"], refpt,
                 float(d[s + "current"]), 2 * tau1, beam.mass, beam.charge, e.type,
                 e.values, e.length, _zedge(d, step), z, rf, ref)
    assert col_err(pk, d[s + "p_kick"]) < 1e-13


@pytest.mark.parametrize("name", ["nl_wake_file", "nl_wake_solrf"])
def test_oracle_kick2wake_trajectory_matches_reference(oracle, name):
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    nsteps = int(d["nsteps"])
    for k, e in enumerate(run.segments(), 1):
        if k > nsteps:
            break
        run.step(e)
        row = run.diag()
        for f, lo in ((24, 5), (25, 11), (26, 17)):
            ref = d[f"fort{f}"][k, 1:]
            got = row[lo:lo + 6]
            scale = np.abs(ref).copy()
            scale[0] = max(scale[0], abs(ref[1]))
            scale[2] = max(scale[2], abs(ref[3]))
            assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
    assert col_err(run.particles(), d["final"]) < 1e-12


@pytest.mark.parametrize("flagmap", [1, 2])
def test_oracle_matches_executable_bends(oracle, flagmap):
    """Bends at zero current through the reference executable
    (tests/golden/exe_bend.npz, both integrators): the moments of every step
    and the particles of the final -2 dump.  Pins Fpol_Dipole, Sector_Dipole
    and Bpol_Dipole with pole-face angles, curvatures, fringe integrals and
    k1, the refptcl(5) advance at the entrance and the z bookkeeping."""
    d = load_golden("exe_bend")
    grid = tuple(int(v) for v in d["grid"])
    lat = fixture_lattice(d)
    run = oracle.Run(d["pts_in"], fixture_beam(d), [e for e in lat if e.type >= 0], grid,
                     flagmap=flagmap)
    k = 0
    for e in lat:
        if e.type < 0:
            continue
        for _ in range(e.nseg):
            run.step(e)
            k += 1
            row = run.diag()
            for f, lo in ((24, 5), (25, 11), (26, 17)):
                ref = d[f"m{flagmap}_fort{f}"][k, 1:]
                got = row[lo:lo + 6]
                scale = np.abs(ref).copy()
                scale[0] = max(scale[0], abs(ref[1]))
                scale[2] = max(scale[2], abs(ref[3]))
                assert np.all(np.abs(got - ref) <= 1e-11 * scale), (k, f, got, ref)
            assert full_row_err(row, fort_row(d, k, f"m{flagmap}_")) < 1e-11, k
            assert run.st.z == pytest.approx(d[f"m{flagmap}_fort24"][k, 0], rel=1e-14)
    assert k == d[f"m{flagmap}_fort24"].shape[0] - 1
    ref = d[f"m{flagmap}_phase60"]
    got = run.particles()
    scale = np.abs(ref).max(axis=0) + 1e-300
    assert np.all(np.abs(got - ref) <= 1e-12 * scale), np.abs(got - ref).max(axis=0)


def test_oracle_isr_closed_forms(oracle):
    """ISRchicane against the textbook synchrotron-radiation results for an
    electron: with rd = 0 the loss over tau is the Larmor energy
    (2/3) r_e gamma^4 tau / rho^2 (in units of m c^2, ultra-relativistic),
    and the spread rd adds has the variance of quantum excitation,
    (55 / (24 sqrt 3)) r_e (hbar / m c) gamma^7 tau / rho^3, in (gamma beta)^2."""
    re, hbar, me, c = 2.81794028946e-15, 1.054571628e-34, 9.109382154e-31, 299792458.0
    mass, gam0, angle, length, tau = 0.511e6, 2000.0, 0.1, 2.0, 0.05
    rho = length / angle
    cst = oracle.physconst(1.3e9)
    v = [angle] + [0.0] * 13
    refpt = np.array([0, 0, 0, 0, 0, -gam0])
    pts = np.zeros((2, 9))
    pts[:, 6] = -1.0 / mass
    pts[:, 7] = 1.0
    oracle.isr_chicane(cst, pts, refpt, tau, length, v, mass, np.array([0.0, 1.0]))
    # in gamma beta; beta^2 = 1 - 2.5e-7 sets the tolerance of the
    # ultra-relativistic forms
    gb0 = np.sqrt(gam0 ** 2 - 1.0)
    gbm = np.sqrt((gam0 - pts[0, 5]) ** 2 - 1.0)
    gb1 = np.sqrt((gam0 - pts[1, 5]) ** 2 - 1.0)
    larmor = 2.0 / 3.0 * re * gam0 ** 4 * tau / rho ** 2
    assert gb0 - gbm == pytest.approx(larmor, rel=2e-6)
    var = 55.0 / (24.0 * np.sqrt(3.0)) * re * hbar / (me * c) * gam0 ** 7 * tau / rho ** 3
    assert (gbm - gb1) ** 2 == pytest.approx(var, rel=4e-6)
    # transverse momenta shrink with the momentum: zero stays zero
    assert np.all(pts[:, [1, 3]] == 0.0)


def test_oracle_matches_executable_charge_states(oracle):
    """Two charge states through the reference executable
    (tests/golden/exe_chg.npz): every column of every diagnostic file, the
    particle counts of fort.28 and the per-state counts of fort.32, which
    chgupdate_BeamBunch recounts after each conv1st as particles of both
    states are lost."""
    d = load_golden("exe_chg")
    grid = tuple(int(v) for v in d["grid"])
    qmcc = d["charge_states"][:, 2]
    run = oracle.Run(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    assert full_row_err(run.diag(), fort_row(d, 0)) < 1e-12
    # before the first step fort.32 holds the deck's nptlist
    assert np.array_equal(d["fort32"][0, 1:], d["charge_states"][:, 0])
    k = 0
    for k, e in enumerate(run.segments(), 1):
        run.step(e)
        assert full_row_err(run.diag(), fort_row(d, k)) < 1e-11, k
        assert d["fort28"][k, 1] == d["fort28"][k, 2] == run.n
        assert d["fort28"][k, 3] == run.st.npt
        counts = oracle.chgupdate(run.particles(), qmcc)
        assert np.array_equal(counts, d["fort32"][k, 1:]), (k, counts)
    assert k == d["fort24"].shape[0] - 1
    assert d["fort32"][-1, 1:].sum() < d["fort32"][0, 1:].sum()   # losses of both states
    assert np.all(d["fort32"][-1, 1:] < d["fort32"][0, 1:])


def test_chgupdate_closed_form(oracle):
    """A particle counts for a state when its q/mc^2 lies within the REAL(4)
    1e-4 of that state, for every state it is that close to."""
    q = 1.0 / 938.272088e6
    eps = float(np.float32(1.0e-4))
    pts = np.zeros((6, 9))
    pts[:, 6] = [q, q * (1 + 0.99 * eps), q * (1 - 1.01 * eps), 2 * q, -q, q]
    assert list(oracle.chgupdate(pts, [q, 2 * q, -q])) == [3, 1, 1]
    assert list(oracle.chgupdate(pts, [q * (1 + 0.5 * eps)])) == [3]
    assert list(oracle.chgupdate(pts, [q * (1 - 0.5 * eps)])) == [3]
    assert list(oracle.chgupdate(pts, [q * (1 + 1.5 * eps)])) == [1]


def test_oracle_csr_wake_matches_reference(oracle):
    """csrwakeTrIGF_FieldQuant restated (or_csr_wake) against the reference's
    own routine on line densities in every regime it distinguishes: the
    entrance transient and steady state inside the bend, lines across the
    exit, the exit transients past it (with the retarded-angle root finder),
    a low gamma, and lines before the entrance (tests/golden/csr.npz,
    written by oracle/_ref/ref_csr)."""
    d = load_golden("csr")
    k = 0
    for nx, r0, ptmin, hx, blen, gam in d["cases"]:
        nx = int(nx)
        rho, ref = d["rho"][k:k + nx], d["ez"][k:k + nx]
        k += nx
        got = oracle.csr_wake(r0, ptmin, hx, blen, rho, gam)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), (r0, ptmin, gam)
        assert got[0] == 0.0


def test_oracle_csr_wake_kind_in_the_loop(oracle):
    """or_wakefield kind 3 bins the particles as the wake does and hands the
    line density to the CSR routine with zwkmin = range(5)/gamma + zoff and
    hzwake/gamma (src/Contrl/AccSimulator.f90:1344-1396): the same as binning
    here and calling the routine directly; no transverse wake."""
    rng = np.random.default_rng(3)
    n, nz, gam = 5000, 32, 50.0
    pts = np.zeros((n, 9))
    pts[:, 0] = rng.normal(0, 1e-3, n)
    pts[:, 2] = rng.normal(0, 1e-3, n)
    pts[:, 4] = rng.normal(0, 2e-4, n)
    pts[:, 6] = 1.0 / 0.511e6
    pts[:, 7] = 1e-9 / n
    zmin, zmax = pts[:, 4].min() * 1.01, pts[:, 4].max() * 1.01
    rg = np.array([-4e-3, 4e-3, -4e-3, 4e-3, zmin, zmax])
    c = oracle.physconst(1.3e9)
    w = oracle.make_wake(3, r0=8.0, zoff=0.3, blength=0.4)
    out = oracle.wakefield(c, pts, rg, nz, 1e-3, n, gam, -1.0, w)
    hz = (zmax - zmin) / (nz - 1)
    dens = np.zeros(nz + 1)
    iz = ((pts[:, 4] - zmin) / hz + 1).astype(int)
    ab = ((zmin - pts[:, 4]) + iz * hz) / hz
    np.add.at(dens, iz - 1, ab)
    np.add.at(dens, iz, 1.0 - ab)
    rho = dens[:nz] * 1e-3 / c.scfreq / n / (hz / gam) * -1.0
    ref = oracle.csr_wake(8.0, zmin / gam + 0.3, hz / gam, 0.4, rho, gam)
    assert np.all(out[0] == 0) and np.all(out[1] == 0)
    assert np.max(np.abs(out[2] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_oracle_csr_switches_follow_the_loop(oracle):
    """The loop's CSR bookkeeping on drift, bend (v3 = 600), drift, quad with
    current: inside the bend the wake's offset is z - zbleng (the bend's
    entrance) and its length the bend's; after it, in every element, the
    offset is measured from the bend's entrance (z - zbleng + sentr) with the
    bend's length; a bend with v3 <= 500 switches the transient off, one with
    v3 <= 200 has no CSR of its own, and no current means no CSR."""
    from impactz.beam import Beam, Element
    beam = Beam(current=1.0, kinetic=100e6, mass=0.51099895e6, charge=-1.0, freq=1.3e9)
    lat = [Element(0.1, 2, 1, 0, 1.0),
           Element(0.4, 4, 1, 4, 0.15, 0.0, 600.0, 0.02),
           Element(0.3, 3, 1, 0, 1.0),
           Element(0.2, 2, 1, 4, 0.1, 0.0, 150.0, 0.02),
           Element(0.1, 1, 1, 0, 1.0)]
    run = oracle.Run(np.zeros((1, 9)), beam, lat, (8, 8, 8))
    z0 = 0.0
    seen = []
    for e in lat:
        run._left, run.zedge = e.nseg, z0
        run._csr_entry(e)
        tau1 = 0.5 * e.length / e.nseg
        for j in range(e.nseg):
            zmid = z0 + (2 * j + 1) * tau1
            w = run.csr_wake(e, zmid)
            seen.append(None if w is None else (w.kind, w.r0, w.zoff, w.blength, zmid))
        run._csr_exit(e)
        z0 += e.length
    r0 = abs(1.0 / (0.15 / 0.4))
    # drift before the bend: no CSR
    assert seen[:2] == [None, None]
    for k in range(2, 6):          # in the bend: offset from its entrance at 0.1
        kind, rr, zoff, blen, zmid = seen[k]
        assert (kind, rr, blen) == (3, r0, 0.4) and zoff == pytest.approx(zmid - 0.1)
    for k in range(6, 9):          # the drift after it: still from the entrance
        kind, rr, zoff, blen, zmid = seen[k]
        assert (kind, rr, blen) == (3, r0, 0.4) and zoff == pytest.approx(zmid - 0.1)
    # a weak bend (v3 = 150): no CSR in it, and its entry switched the
    # transient off for the drift after it
    assert seen[9:] == [None, None, None]
