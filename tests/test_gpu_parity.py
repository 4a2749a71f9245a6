"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

Bar (FP64 form): 1e-10 relative, each quantity normalised by its own scale
(max |rho|, max |phi|, per-coordinate max |Pts1 row|, rms for centroids).
The mixed form is held to the tolerances stated per test.
"""
import json
import os

import numpy as np
import pytest

from conftest import (ROOT, col_err, element_entrance, fixture_beam, fixture_lattice,
                      fixture_row_err, full_row_err, load_golden, rel_err, segment)

pytestmark = pytest.mark.gpu

TOL = 1e-10
LOG = os.path.join(ROOT, "gpurun_out", "parity_errors.jsonl")


def record(**kw):
    """Append measured errors (for the record; assertions do the checking)."""
    try:
        os.makedirs(os.path.dirname(LOG), exist_ok=True)
        with open(LOG, "a") as f:
            f.write(json.dumps(kw) + "\n")
    except OSError:
        pass
    print(kw, flush=True)


CASES = [("fodo_16cube", 1), ("fodo_16cube", 5), ("fodo_aniso", 2),
         ("fodo_lossy", 1), ("fodo_lossy", 2), ("cf_sol", 2), ("cf_sol", 10),
         ("kquad", 3), ("rf", 6), ("rf", 22), ("rf", 30)]


@pytest.mark.parametrize("name,step", CASES)
def test_gpu_stages_match_reference(gpu, name, step):
    hip = gpu
    import ctypes
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    tau1 = 0.5 * e.length / e.nseg
    b = hip.BeamBunch.from_beam(d[s + "p0"], beam, hip.FP64)
    info = b.info()
    for k in range(6):
        info.refpt[k] = float(d[s + "refpt0"][k])
    info.current = float(d[s + "current0"])
    hip.check(hip.lib().izh_bunch_set_info(b.h, ctypes.byref(info)))
    field = hip.FieldQuant(*grid)
    dom = hip.CompDom(*grid)
    rf, zedge = hip.RFData.for_element(e), element_entrance(d, step)
    print(f"stages {name}:{step}", flush=True)
    err = {}
    z = b.map1(float(d[s + "z0"]), tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
    err["map1a"] = col_err(b.getpts(), d[s + "p_map1a"])
    pr, nplc, nptot = b.conv1st(2 * tau1, e.radius, e.radius)
    assert nplc == nptot == d[s + "p_conv1st"].shape[0]
    err["ptrange"] = rel_err(pr, d[s + "ptrange"])
    err["current"] = abs(b.current / float(d[s + "current"]) - 1)
    err["conv1st"] = col_err(b.getpts(), d[s + "p_conv1st"])
    dom.update(pr)
    err["range"] = rel_err(dom.range, d[s + "range"])
    err["msize"] = rel_err(dom.msize, d[s + "msize"])
    b.charge(dom, field)
    err["rho"] = rel_err(field.download(hip.RHO), d[s + "rho"])
    field.update3O(dom)
    err["phi"] = rel_err(field.download(hip.PHI), d[s + "phi"])
    b.cvbkforth1st()
    err["cvbk"] = col_err(b.getpts(), d[s + "p_cvbk"])
    b.kick1(2 * tau1, field, dom)
    err["kick"] = col_err(b.getpts(), d[s + "p_kick"])
    z = b.map1(z, tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
    err["map1b"] = col_err(b.getpts(), d[s + "p_map1b"])
    err["refpt"] = rel_err(b.refpt, d[s + "refpt"])
    assert z == pytest.approx(float(d[s + "z"]), rel=1e-15)
    record(test="stages", case=f"{name}:{step}", **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


def test_gpu_solver_matches_direct_convolution(gpu, oracle):
    hip = gpu
    rng = np.random.default_rng(11)
    for grid, h in (((8, 8, 16), (0.3, 0.5, 0.9)), ((16, 8, 8), (1.0, 0.2, 0.4))):
        nx, ny, nz = grid
        rho = rng.standard_normal((nz, ny, nx))
        f = hip.FieldQuant(*grid)
        dom = hip.CompDom(*grid)
        dom.set([0, (nx - 1) * h[0], 0, (ny - 1) * h[1], 0, (nz - 1) * h[2]], h)
        f.upload_rho(rho)
        f.update3O(dom)
        phi = f.download(hip.PHI)
        ref = oracle.solve_direct(rho, h)
        e = rel_err(phi, ref)
        record(test="solver_direct", grid=grid, err=e)
        assert e < 1e-12


@pytest.mark.parametrize("name", ["fodo_traj", "fodo_16cube", "fodo_lossy", "cf_sol",
                                  "kquad", "rf"])
def test_gpu_trajectory_matches_reference(gpu, name):
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    rows = sim.run(nsteps=int(d["nsteps"]))
    worst = 0.0
    for k in range(rows.shape[0]):
        worst = max(worst, fixture_row_err(rows[k, 1:], d, k))
    fin = sim.bunch.getpts()
    assert fin.shape == d["final"].shape
    pe = col_err(fin, d["final"])
    ce = abs(sim.bunch.current / float(d["final_current"]) - 1)
    record(test="trajectory", case=name, steps=int(d["nsteps"]), moments=worst,
           particles=pe, current=ce)
    assert worst < TOL and pe < TOL and ce < 1e-14


def test_gpu_matches_oracle_100k_32cube(gpu, oracle):
    """C1 sizes (100k particles, 32^3) for 8 steps of the FODO lattice."""
    from impactz import beam as B
    from impactz.sim import Simulation
    bm = B.Beam()
    pts = B.gaussian_bunch(100_000, seed=1)
    lat = B.fodo(1)
    sim = Simulation(pts, bm, lat, (32, 32, 32))
    run = oracle.Run(pts, bm, lat, (32, 32, 32))
    worst = {"rho": 0.0, "phi": 0.0, "moments": 0.0, "pts": 0.0}
    for k, e in enumerate(sim.segments(), 1):
        if k > 8:
            break
        mom = sim.step(e)
        rho_o, phi_o = run.step(e, want_fields=True)
        worst["rho"] = max(worst["rho"], rel_err(sim.field.download(gpu.RHO), rho_o))
        worst["phi"] = max(worst["phi"], rel_err(sim.field.download(gpu.PHI), phi_o))
        worst["moments"] = max(worst["moments"], full_row_err(sim.diag(mom), run.diag()))
    worst["pts"] = col_err(sim.bunch.getpts(), run.particles())
    record(test="oracle_100k_32cube", **worst)
    for k, v in worst.items():
        assert v < TOL, (k, v)


def test_gpu_deterministic_and_granular_equals_fused(gpu):
    from impactz import beam as B
    from impactz.sim import Simulation
    bm = B.Beam()
    pts = B.gaussian_bunch(50_000, seed=4)
    lat = B.fodo(1)
    outs = []
    for mode in ("step", "step", "granular"):
        sim = Simulation(pts, bm, lat, (16, 16, 32))
        for k, e in enumerate(sim.segments()):
            if k >= 6:
                break
            if mode == "step":
                sim.step(e, want_moments=False)
            else:
                sim.step_granular(e)
        outs.append((sim.bunch.getpts(), sim.field.download(gpu.RHO),
                     sim.field.download(gpu.PHI)))
        sim.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    for a, b in zip(outs[0], outs[2]):
        assert np.array_equal(a, b)


def test_gpu_mixed_precision_error(gpu):
    """Mixed form vs FP64 form on the same input: field and rms drift."""
    from impactz import beam as B
    from impactz.sim import Simulation
    bm = B.Beam()
    pts = B.gaussian_bunch(100_000, seed=1)
    lat = B.fodo(2)
    sims = [Simulation(pts, bm, lat, (32, 32, 32), precision=p) for p in (gpu.FP64, gpu.MIXED)]
    field_err = 0.0
    for k, e in enumerate(sims[0].segments()):
        moms = [s.step(e) for s in sims]
        ef = [np.stack([s.field.download(w) for w in (gpu.EX, gpu.EY, gpu.EZ)]) for s in sims]
        field_err = max(field_err, float(np.linalg.norm(ef[1] - ef[0]) / np.linalg.norm(ef[0])))
    r64, r32 = (s.diag() for s in sims)
    rms_cols = [6, 8, 12, 14, 18, 20]
    rms_err = float(np.max(np.abs(r32[rms_cols] - r64[rms_cols]) / np.abs(r64[rms_cols])))
    record(test="mixed_vs_fp64", steps=32, field_l2=field_err, rms=rms_err)
    assert field_err < 1e-4
    assert rms_err < 1e-4


def test_gpu_edge_cases(gpu):
    import ctypes
    from impactz import beam as B
    bm = B.Beam()
    hip = gpu
    # every particle lost in the first conv1st: empty bunch, zero current,
    # and later steps still run
    pts = B.gaussian_bunch(2000, seed=2)
    b = hip.BeamBunch.from_beam(pts, bm)
    f = hip.FieldQuant(8, 8, 8)
    dom = hip.CompDom(8, 8, 8)
    z = b.drift1(0.0, 0.01, 0)
    pr, nplc, nptot = b.conv1st(0.02, 1e-9, 1e-9)
    assert nplc == 0 and nptot == 0 and b.current == 0.0
    assert b.getpts().shape == (0, 9)
    b.cvbkforth1st()
    b.kick1(0.02, f, dom)          # zero current: no kick, back to the z frame
    # an empty bunch can be stepped and gives zero moments
    z, mom = b.step(f, dom, z, hip.segment(0.05, 0, 1.0, 1, 1.0, 1.0), bm.current, True)
    assert np.all(mom == 0)
    # a single particle has a degenerate domain: reported, not computed
    one = hip.BeamBunch.from_beam(pts[:1], bm)
    with pytest.raises(hip.IZHError, match="mesh size"):
        one.step(f, dom, 0.0, hip.segment(0.05, 0, 1.0, 1, 1.0, 1.0), bm.current)
    # stage order is enforced
    c = hip.BeamBunch.from_beam(pts, bm)
    with pytest.raises(hip.IZHError, match="conv1st"):
        c.charge(dom, f)
    with pytest.raises(hip.IZHError, match="z frame"):
        c.conv1st(0.02, 1.0, 1.0)
        c.drift1(0.0, 0.01, 0)
    del ctypes


def test_gpu_unsupported_elements_are_refused(gpu):
    """drift1_BeamBunch branches off this path fail with a message, before
    touching the particles."""
    from impactz import beam as B
    hip = gpu
    bm = B.Beam()
    pts = B.gaussian_bunch(1000, seed=2)
    b = hip.BeamBunch.from_beam(pts, bm)
    for bt, v, what in ((1, [1.0, -20.0, 1.0], "undulator"),
                        (1, [1.0, 2.0, 1.0], "file"),
                        (2, [1.0, 1.0, 1.0, -1.0], "ideal RF"),
                        (4, [0.0], "not on this path")):
        with pytest.raises(hip.IZHError, match=what):
            b.map1(0.0, 0.01, bt, v, 10)
    assert col_err(b.getpts(), pts) == 0.0
