"""The nonlinear Lorentz integrator (Flagmap=2) on the GPU against the
compiled reference (tests/golden/nl_*.npz): every stage of chosen steps
(drift2, the space-charge stages, kick2 with the element's field, drift2)
and the fort.24-26 trajectory, to 1e-10 relative."""
import ctypes

import numpy as np
import pytest

from conftest import col_err, diag_row_err, element_entrance, fixture_beam, \
    fixture_lattice, load_golden, rel_err, segment
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu

CASES = [("nl_fodo", 1), ("nl_fodo", 5), ("nl_cf_sol", 2), ("nl_cf_sol", 10),
         ("nl_cf_sol", 11)]


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
    b.kick2(z, 2 * tau1, hip.element(e.type, e.values, e.length, element_entrance(d, step)), field,
            dom)
    err["kick2"] = col_err(b.getpts(), d[s + "p_kick"])
    z = b.drift2(z, 2 * tau1)
    err["drift2b"] = col_err(b.getpts(), d[s + "p_map1b"])
    err["refpt"] = rel_err(b.refpt, d[s + "refpt"])
    assert z == pytest.approx(float(d[s + "z"]), rel=1e-15)
    record(test="nonlinear_stages", case=f"{name}:{step}", **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("name", ["nl_fodo", "nl_cf_sol"])
def test_gpu_nonlinear_trajectory_matches_reference(gpu, name):
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    print(f"nonlinear trajectory {name}", flush=True)
    rows = sim.run(nsteps=int(d["nsteps"]))
    worst = 0.0
    for k in range(1, rows.shape[0]):
        ref = np.concatenate([d["fort24"][k, 1:], d["fort25"][k, 1:], d["fort26"][k, 1:]])
        worst = max(worst, diag_row_err(rows[k, 1:], ref, ref_is_fort=True))
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
