import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "impact-z_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def fixture_beam(d):
    from impactz.beam import Beam
    cur, kin, mass, chg, freq, ph = (float(v) for v in d["beam"])
    return Beam(current=cur, kinetic=kin, mass=mass, charge=chg, freq=freq, phase0=ph)


def fixture_lattice(d, key="lattice"):
    """The lattice of a fixture: rows of length, nseg, mapstp, type and the
    values, with RF fields stored as field<i>."""
    from impactz.beam import Element
    out = []
    for i, r in enumerate(d[key]):
        v = [float(x) for x in r[4:]]
        fk = f"field{i}"
        out.append(Element(float(r[0]), int(r[1]), int(r[2]), int(r[3]), *v[:4],
                           more=tuple(v[4:]), field=d[fk] if fk in d else None))
    return out


def element_entrance(d, step):
    """z of the entrance of the element the 1-based step belongs to."""
    z, k = 0.0, 0
    for e in fixture_lattice(d):
        if k + e.nseg >= step:
            return z
        k += e.nseg
        z += e.length
    raise IndexError(step)


def segment(d, step):
    """Element of the 1-based step of a fixture's lattice."""
    lat = fixture_lattice(d)
    k = 0
    for e in lat:
        for _ in range(e.nseg):
            k += 1
            if k == step:
                return e
    raise IndexError(step)


def rel_err(a, b):
    """max |a-b| / max |b|: error normalised by the quantity's own scale."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a - b)))


def col_err(a, b):
    """Per-column (per Pts1 row) max error normalised by that column's scale."""
    a = np.asarray(a)
    b = np.asarray(b)
    return max(rel_err(a[:, c], b[:, c]) for c in range(b.shape[1]))


def diag_row_err(got, ref, ref_is_fort=False):
    """Worst error of the fort.24/25/26 columns of a diagnostic row.

    ``got`` is a 41-value row (impactz.sim.diag_row order); ``ref`` is either
    such a row or, with ref_is_fort, the 18 fort.24+25+26 values.  Each
    column is normalised by its own magnitude, centroids by the rms size.
    """
    g = np.asarray(got)[5:23]
    r = np.asarray(ref)[:18] if ref_is_fort else np.asarray(ref)[5:23]
    worst = 0.0
    for b in range(3):
        rb, gb = r[6 * b:6 * b + 6], g[6 * b:6 * b + 6]
        scale = np.abs(rb).copy()
        scale[0] = max(scale[0], abs(rb[1]))
        scale[2] = max(scale[2], abs(rb[3]))
        worst = max(worst, float(np.max(np.abs(gb - rb) / scale)))
    return worst


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, initialised on device 0 (GPU tests only)."""
    from impactz import hip
    hip.init(0)
    return hip
