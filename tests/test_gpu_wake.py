"""1D wakefields on the GPU against the compiled reference: the wake fields
of izh_wakefield and kick1wake stage by stage (tests/golden/wake_*.npz,
written by the reference's wakefield_FieldQuant / wakefieldread_FieldQuant
and kick1wake_BeamBunch), and whole runs step by step, through izh_run and
through the deck against the reference's fort.24-26 (including the
executable's own -41 run, tests/golden/exe_wake.npz)."""
import numpy as np
import pytest

from conftest import (col_err, deck_files_err, fixture_row_err, element_entrance, fixture_beam,
                      fixture_lattice, load_golden, segment)
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu

# (fixture, dumped step, wake kind, scale, flagbtw)
CASES = [("wake_file", 2, 2, 1.5, 0), ("wake_file", 12, 2, 0.8, 2),
         ("wake_solrf", 6, 1, 1.0, 0)]


@pytest.mark.parametrize("name,step,kind,scale,btw", CASES)
def test_gpu_wake_stages_match_reference(gpu, name, step, kind, scale, btw):
    import ctypes
    hip = gpu
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    tau1 = 0.5 * e.length / e.nseg
    lat = fixture_lattice(d)
    if kind == 2:
        table = hip.RFData([x for x in lat if x.type == -41][0].field, fourier=False)
        w = hip.wake(2, table=table, flagbtw=btw, scale=scale)
    else:
        v = e.values
        w = hip.wake(1, v[11], v[12], v[13], scale=scale)
    b = hip.BeamBunch.from_beam(d[s + "p0"], beam, hip.FP64)
    info = b.info()
    for k in range(6):
        info.refpt[k] = float(d[s + "refpt0"][k])
    info.current = float(d[s + "current0"])
    hip.check(hip.lib().izh_bunch_set_info(b.h, ctypes.byref(info)))
    field = hip.FieldQuant(*grid)
    dom = hip.CompDom(*grid)
    rf, zedge = hip.RFData.for_element(e), element_entrance(d, step)
    print(f"wake stages {name}:{step}", flush=True)
    err = {}
    z = b.map1(float(d[s + "z0"]), tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
    b.conv1st(2 * tau1, e.radius, e.radius)
    dom.update(d[s + "ptrange"])
    b.charge(dom, field)
    field.update3O(dom)
    b.wakefield(field, dom, w)
    got, ref = field.wake_download(), d[s + "wake"]
    for k, nm in enumerate(("exwake", "eywake", "ezwake")):
        err[nm] = float(np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300))
    b.cvbkforth1st()
    b.kick1wake(2 * tau1, field, dom)
    err["kick"] = col_err(b.getpts(), d[s + "p_kick"])
    z = b.map1(z, tau1, e.type, e.values, e.mapstp, e.length, zedge, rf)
    err["map1b"] = col_err(b.getpts(), d[s + "p_map1b"])
    record(test="wake_stages", case=f"{name}:{step}", **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("name,fused", [("wake_file", False), ("wake_file", True),
                                        ("wake_solrf", False), ("wake_solrf", True),
                                        ("exe_wake", True)])
def test_gpu_wake_runs_match_reference(gpu, name, fused):
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid)
    nsteps = d["fort24"].shape[0] - 1
    print(f"wake run {name} fused={fused}", flush=True)
    if fused:
        recs = sim.run_fused(nsteps)
        rows = [sim.diag(r) for r in recs]
    else:
        rows = [sim.diag(sim.step(e)) for _, e in zip(range(nsteps), sim.segments())]
    worst = 0.0
    for k in range(nsteps):
        worst = max(worst, fixture_row_err(rows[k], d, k + 1))
    record(test="wake_run", case=name, fused=fused, moments=worst)
    assert worst < TOL
    if "final" in d:
        assert col_err(sim.bunch.getpts(), d["final"]) < TOL


@pytest.mark.parametrize("fused", [False, True])
def test_gpu_wake_deck_matches_reference_executable(tmp_path, gpu, fused):
    """The -41 deck the reference executable ran, through impactz.deck."""
    from impactz.beam import write_deck, write_particle_in
    from impactz.deck import run_deck
    d = load_golden("exe_wake")
    grid = tuple(int(v) for v in d["grid"])
    write_deck(str(tmp_path / "ImpactZ.in"), d["pts_in"].shape[0], grid, fixture_beam(d),
               fixture_lattice(d))
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    print(f"wake deck fused={fused}", flush=True)
    run_deck(str(tmp_path), fused=fused)
    worst = deck_files_err(str(tmp_path), d)
    record(test="wake_deck", case="exe_wake", fused=fused, moments=worst)
    assert worst < TOL


def test_gpu_solrf_wake_needs_its_wake(gpu):
    from impactz import beam as B
    hip = gpu
    bm = B.Beam()
    b = hip.BeamBunch.from_beam(B.gaussian_bunch(1000, seed=2), bm)
    f = hip.FieldQuant(8, 8, 8)
    dom = hip.CompDom(8, 8, 8)
    rf = hip.RFData(np.array([3.0, 0.0, 0.2, 0.2, 0.4, -1.0, 0.05]), fourier=True)
    v = [3e6, 324e6, 45.0, 2.0, 1.0, 0, 0, 0, 0, 0, 0.4, 0.01, 0.03, 0.033]
    seg = hip.segment(0.01, 105, v, 10, 1.0, 1.0, 0.2, 0.0, rf)
    with pytest.raises(hip.IZHError, match="needs its izh_wake"):
        b.step(f, dom, 0.0, seg, bm.current)
    with pytest.raises(hip.IZHError, match="needs its izh_wake"):
        b.run(f, dom, 0.0, [seg], bm.current)


@pytest.mark.parametrize("name,step,kind", [("nl_wake_file", 6, 2), ("nl_wake_solrf", 7, 1)])
def test_gpu_kick2wake_stages_match_reference(gpu, name, step, kind):
    """Flagmap=2 with a wake: izh_wakefield after drift2 and izh_kick2wake."""
    import ctypes
    hip = gpu
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    beam = fixture_beam(d)
    e = segment(d, step)
    s = f"s{step}_"
    tau1 = 0.5 * e.length / e.nseg
    lat = fixture_lattice(d)
    if kind == 2:
        table = hip.RFData([x for x in lat if x.type == -41][0].field, fourier=False)
        w = hip.wake(2, table=table, scale=1.5)
    else:
        v = e.values
        w = hip.wake(1, v[11], v[12], v[13])
    b = hip.BeamBunch.from_beam(d[s + "p0"], beam, hip.FP64)
    info = b.info()
    for k in range(6):
        info.refpt[k] = float(d[s + "refpt0"][k])
    info.current = float(d[s + "current0"])
    hip.check(hip.lib().izh_bunch_set_info(b.h, ctypes.byref(info)))
    field = hip.FieldQuant(*grid)
    dom = hip.CompDom(*grid)
    rf = None if e.field is None else hip.RFData(e.field, fourier=True)
    elem = hip.element(e.type, e.values, e.length, element_entrance(d, step), rf=rf)
    print(f"kick2wake stages {name}:{step}", flush=True)
    err = {}
    z = b.drift2(float(d[s + "z0"]), 2 * tau1)
    b.conv1st(2 * tau1, e.radius, e.radius)
    dom.update(d[s + "ptrange"])
    b.charge(dom, field)
    field.update3O(dom)
    b.wakefield(field, dom, w)
    got, ref = field.wake_download(), d[s + "wake"]
    for k, nm in enumerate(("exwake", "eywake", "ezwake")):
        err[nm] = float(np.abs(got[k] - ref[k]).max() / max(np.abs(ref[k]).max(), 1e-300))
    b.cvbkforth1st()
    b.kick2wake(z, 2 * tau1, elem, field, dom)
    err["kick2wake"] = col_err(b.getpts(), d[s + "p_kick"])
    b.drift2(z, 2 * tau1)
    err["drift2b"] = col_err(b.getpts(), d[s + "p_map1b"])
    record(test="kick2wake_stages", case=f"{name}:{step}", **err)
    for k, v in err.items():
        assert v < TOL, (k, v)


@pytest.mark.parametrize("name", ["nl_wake_file", "nl_wake_solrf"])
def test_gpu_kick2wake_runs_match_reference(gpu, name):
    from impactz.sim import Simulation
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    sim = Simulation(d["pts_in"], fixture_beam(d), fixture_lattice(d), grid, flagmap=2)
    nsteps = int(d["nsteps"])
    print(f"kick2wake run {name}", flush=True)
    rows = [sim.diag(sim.step(e)) for _, e in zip(range(nsteps), sim.segments())]
    worst = 0.0
    for k in range(nsteps):
        worst = max(worst, fixture_row_err(rows[k], d, k + 1))
    pe = col_err(sim.bunch.getpts(), d["final"])
    record(test="kick2wake_run", case=name, moments=worst, particles=pe)
    assert worst < TOL and pe < TOL
