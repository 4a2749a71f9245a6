"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-only entry points (physical constants,
quadrupole map, domain update, argument validation) match the reference."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fixture_beam, fixture_lattice, load_golden, segment


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "impactz_hip.h")).read()
    return sorted(set(re.findall(r"\b(izh_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from impactz import hip
    lib = hip.lib()
    declared = header_symbols()
    assert len(declared) >= 20
    for sym in declared:
        assert hasattr(lib, sym), sym
    assert sorted(hip.SYMBOLS) == declared


def test_fortran_module_binds_every_declared_symbol():
    """impact-z_amd/fortran/impactz_hip.f90 has a bind(C) interface for every
    entry point of the header, and nothing else under that prefix."""
    src = open(os.path.join(ROOT, "impact-z_amd", "fortran", "impactz_hip.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(izh_[A-Za-z0-9_]+)"\)', src)))
    assert bound == header_symbols()


def test_fortran_host_is_built():
    from impactz import _build
    assert os.access(_build.FORTRAN_EXE, os.X_OK)


def test_physconst_matches_oracle(oracle):
    from impactz import hip
    c = hip.physconst(324e6)
    o = oracle.physconst(324e6)
    assert (c.clight, c.pi, c.scxl, c.scfreq) == (o.clight, o.pi, o.scxl, o.scfreq)


@pytest.mark.parametrize("name,step", [("fodo_16cube", 1), ("fodo_aniso", 2)])
def test_quad_map_matches_reference(oracle, name, step):
    """Host RK6 map of the library vs the oracle, and the reference's refpt."""
    from impactz import hip
    d = load_golden(name)
    beam = fixture_beam(d)
    e = segment(d, step)
    tau1 = 0.5 * e.length / e.nseg
    c = hip.physconst(beam.freq)
    refpt = d[f"s{step}_refpt0"]
    xm, r1 = hip.quad_maplinear(c, tau1, e.v1, e.mapstp, refpt, beam.charge, beam.mass)
    xo, r2 = oracle.quad_map(oracle.physconst(beam.freq), tau1, e.v1, e.mapstp,
                             refpt, beam.charge, beam.mass)
    assert np.allclose(xm, xo, rtol=1e-14, atol=1e-18)
    assert r1[4] == pytest.approx(r2[4], rel=1e-15)
    # two half steps of the quad move the reference phase as the reference did
    _, r3 = hip.quad_maplinear(c, tau1, e.v1, e.mapstp, r1, beam.charge, beam.mass)
    assert r3[4] == pytest.approx(float(d[f"s{step}_refpt"][4]), rel=1e-14)


@pytest.mark.parametrize("bitype,v", [(1, [10.0, 0.0, 1.0, 0.0]), (1, [-7.5, 0.0, 1.0, 0.0]),
                                      (2, [30.0, 20.0, 5.0, 1.0]), (2, [-4.0, 9.0, 0.0, 1.0]),
                                      (3, [1.5, 0.0, 1.0, 0.0]), (3, [-0.8, 0.0, 1.0, 0.0])])
def test_element_maplinear_matches_oracle(oracle, bitype, v):
    """Host element maps of the library (quadrupole, constant focusing,
    solenoid) against the oracle, itself pinned to the compiled reference."""
    from impactz import beam as B
    from impactz import hip
    bm = B.Beam()
    refpt = np.zeros(6)
    refpt[5] = -bm.gamma0
    xm, r1 = hip.element_maplinear(hip.physconst(bm.freq), 0.0125, bitype, v, 10, refpt,
                                   bm.charge, bm.mass)
    xo, r2 = oracle.element_map(oracle.physconst(bm.freq), 0.0125, bitype, v, 10, refpt,
                                bm.charge, bm.mass)
    assert np.allclose(xm, xo, rtol=1e-14, atol=1e-18)
    assert np.array_equal(xm == 0.0, xo == 0.0)
    assert r1[4] == pytest.approx(r2[4], rel=1e-15)


@pytest.mark.parametrize("idx", [1, 3, 4, 5, 7])
@pytest.mark.parametrize("z_in", [0.0, 0.031, 0.2])
def test_rf_maplinear_matches_oracle(oracle, idx, z_in):
    """RF cavity maps of the library (DTL, CCDTL, CCL, SC on the field table;
    SolRF on the Fourier series) against the oracle, at the entrance, inside
    the DTL's first quadrupole and past the field's end."""
    from impactz import hip
    from conftest import fixture_beam
    d = load_golden("rf")
    bm = fixture_beam(d)
    e = fixture_lattice(d)[idx]
    zedge = 1.0
    refpt = np.zeros(6)
    refpt[4], refpt[5] = 2.3, -bm.gamma0
    tau = 0.5 * e.length / e.nseg
    xm, r1 = hip.element_maplinear(hip.physconst(bm.freq), tau, e.type, e.values, e.mapstp,
                                   refpt, bm.charge, bm.mass, zedge + z_in, e.length, zedge,
                                   hip.RFData.for_element(e))
    xo, r2 = oracle.element_map(oracle.physconst(bm.freq), tau, e.type, e.values, e.mapstp,
                                refpt, bm.charge, bm.mass, zedge + z_in, e.length, zedge,
                                oracle.RFData(e.field, fourier=e.type == 105))
    assert np.allclose(xm, xo, rtol=1e-13, atol=1e-16)
    assert np.array_equal(xm == 0.0, xo == 0.0)
    assert np.allclose(r1, r2, rtol=1e-15)


def test_rfdata_is_checked():
    from impactz import hip
    tab = np.array([[0.0, 1.0, 0.0, 0.0], [1.0, 1.0, 0.0, 0.0]])
    c = hip.physconst(324e6)
    r = np.zeros(6)
    r[5] = -1.16
    v = [1e6, 324e6, 0.0, 1.0, 1.0]
    with pytest.raises(hip.IZHError, match="announces"):
        hip.element_maplinear(c, 0.01, 105, v, 10, r, 1.0, 938e6, length=1.0,
                              rf=hip.RFData(np.array([9.0, 0.0, 1.0, 1.0, 0.5, 0.2]),
                                            fourier=True))
    with pytest.raises(hip.IZHError, match="needs its field"):
        hip.element_maplinear(c, 0.01, 103, v, 10, r, 1.0, 938e6, length=1.0)
    with pytest.raises(hip.IZHError, match="Fourier"):
        hip.element_maplinear(c, 0.01, 105, v, 10, r, 1.0, 938e6, length=1.0,
                              rf=hip.RFData(tab, fourier=False))


def test_element_maplinear_refuses_other_elements():
    from impactz import hip
    c = hip.physconst(324e6)
    r = np.zeros(6)
    r[5] = -1.16
    with pytest.raises(hip.IZHError, match="file"):
        hip.element_maplinear(c, 0.01, 1, [1.0, 1.0, 0.0, 0.0], 10, r, 1.0, 938e6)
    with pytest.raises(hip.IZHError, match="no linear map"):
        hip.element_maplinear(c, 0.01, 4, [1.0, 0.0, 0.0, 0.0], 10, r, 1.0, 938e6)
    with pytest.raises(hip.IZHError, match="no linear map"):
        hip.element_maplinear(c, 0.01, 106, [1.0, 324e6, 0.0, 1.0], 10, r, 1.0, 938e6)


@pytest.mark.parametrize("name,step", [("fodo_16cube", 1), ("fodo_16cube", 5),
                                       ("fodo_aniso", 2), ("fodo_lossy", 2)])
def test_update_compdom_matches_reference(name, step):
    from impactz import hip
    d = load_golden(name)
    grid = tuple(int(v) for v in d["grid"])
    dom = hip.CompDom(*grid)
    dom.update(d[f"s{step}_ptrange"])
    assert np.array_equal(dom.range, d[f"s{step}_range"])
    assert np.array_equal(dom.msize, d[f"s{step}_msize"])


def test_argument_errors_are_reported():
    from impactz import hip
    dom = hip.CompDom(3, 8, 8)
    with pytest.raises(hip.IZHError, match="below 4"):
        dom.update(np.array([-1, 1, -1, 1, -1, 1.0]))
    dom = hip.CompDom(8, 8, 8)
    with pytest.raises(hip.IZHError, match="Flagbc=1"):
        dom.update(np.array([-1, 1, -1, 1, -1, 1.0]), flagbc=2)
    with pytest.raises(hip.IZHError, match="powers of two"):
        hip.FieldQuant(12, 8, 8)
    c = hip.physconst(324e6)
    with pytest.raises(hip.IZHError, match="bad argument"):
        hip.quad_maplinear(c, 0.1, 1.0, 0, np.zeros(6), 1.0, 938e6)
