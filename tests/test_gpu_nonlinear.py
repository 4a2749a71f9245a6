"""The nonlinear Lorentz integrator (Flagmap=2) on the GPU against the
compiled reference (tests/golden/nl_*.npz): every stage of chosen steps
(drift2, the space-charge stages, kick2 with the element's field, drift2)
and the fort.24-26 trajectory, to 1e-10 relative."""
import ctypes

import numpy as np
import pytest

from conftest import col_err, deck_files_err, element_entrance, fixture_beam, fixture_row_err, \
    fixture_lattice, load_golden, rel_err, segment
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu

CASES = [("nl_fodo", 1), ("nl_fodo", 5), ("nl_cf_sol", 2), ("nl_cf_sol", 10),
         ("nl_cf_sol", 11), ("nl_rf", 6), ("nl_rf", 22), ("nl_solrf", 7)]


@pytest.mark.parametrize("name,step", CASES)
def test_gpu_nonlinear_stages_match_reference(gpu, name, step):
    hip = gpu
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
    print(f"nonlinear stages {name}:{step}", flush=True)
    err = {}
    z = b.drift2(float(d[s + "z0"]), 2 * tau1)
    err["drift2a"] = col_err(b.getpts(), d[s + "p_map1a"])
    pr, nplc, nptot = b.conv1st(2 * tau1, e.radius, e.radius)
    err["conv1st"] = col_err(b.getpts(), d[s + "p_conv1st"])
    dom.update(pr)
    b.charge(dom, field)
    field.update3O(dom)
    err["phi"] = rel_err(field.download(hip.PHI), d[s + "phi"])
    b.cvbkforth1st()
    rf = None if e.field is None else hip.RFData(e.field, fourier=True)
    b.kick2(z, 2 * tau1,
            hip.element(e.type, e.values, e.length, element_entrance(d, step), rf=rf), field,
            dom)
    err["kick2"] = col_err(b.getpts(), d[s + "p_kick"])
    z = b.drift2(z, 2 * tau1)
    err["drift2b"] = col_err(b.getpts(), d[s + "p_map1b"])
    err["refpt"] = rel_err(b.refpt, d[s + "refpt"])
    assert z == pytest.approx(float(d[s + "z"]), rel=1e-15)
    record(test="nonlinear_stages", case=f"{name}:{step}", **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("name", ["nl_fodo", "nl_cf_sol", "nl_rf", "nl_solrf"])
def test_gpu_nonlinear_trajectory_matches_reference(gpu, name):
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    print(f"nonlinear trajectory {name}", flush=True)
    rows = sim.run(nsteps=int(d["nsteps"]))
    worst = 0.0
    for k in range(rows.shape[0]):
        worst = max(worst, fixture_row_err(rows[k, 1:], d, k))
    pe = col_err(sim.bunch.getpts(), d["final"])
    record(test="nonlinear_trajectory", case=name, moments=worst, particles=pe)
    assert worst < TOL and pe < TOL
    with pytest.raises(gpu.IZHError, match="Flagmap=2"):
        sim.run_fused(2)


def test_gpu_kick2_without_space_charge_and_refusals(gpu):
    """scatter20 on a z-frame bunch: a drift's (zero) field leaves it alone;
    elements without an analytic field are refused."""
    from impactz import beam as B
    hip = gpu
    bm = B.Beam()
    pts = B.gaussian_bunch(2000, seed=5)
    b = hip.BeamBunch.from_beam(pts, bm)
    b.kick2(0.0, 0.01, hip.element(0, [1.0], 0.1, 0.0), None, None)
    assert col_err(b.getpts(), pts) == 0.0
    for bt, v, what in ((1, [10.0, 1.0], "file"), (5, [2.0], "not on this path")):
        with pytest.raises(hip.IZHError, match=what):
            b.kick2(0.0, 0.01, hip.element(bt, v, 0.1, 0.0), None, None)


@pytest.mark.parametrize("fused", [False, True])
def test_gpu_reference_example1_deck(tmp_path, gpu, fused):
    """The reference's examples/Example1 (H- through nine DTL cells on their
    Fourier fields, Lorentz integrator, 1002 steps) as a deck through
    impactz.deck, against the fort.24-26 its executable wrote
    (tests/golden/example1.npz).  fused only fuses Flagmap=1 stretches, so
    both settings run the steps one by one here."""
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden("example1")
    grid = tuple(int(v) for v in d["grid"])
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid, fixture_beam(d),
               fixture_lattice(d), flagmap=2)
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    print(f"example1 deck fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    f = {u: np.loadtxt(tmp_path / f"fort.{u}", ndmin=2) for u in (18, 24)}
    assert f[24].shape == d["fort24"].shape
    worst = deck_files_err(str(tmp_path), d)
    gam = np.abs(f[18][:, 2] / d["fort18"][:, 2] - 1).max()
    record(test="example1_deck", fused=fused, steps=f[24].shape[0] - 1, moments=worst,
           gamma=float(gam))
    assert worst < TOL and gam < 1e-13
