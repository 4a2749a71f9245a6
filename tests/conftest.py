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


# Where each fort file's columns (after z) sit in a 41-value diagnostic row:
# fort.18 refpt(5), gamma, energy, beta, r_max; fort.24/25/26 centroid, rms,
# momentum centroid, momentum rms, alpha, emittance; fort.27 the six maxima;
# fort.29 the cube roots of the third central moments; fort.30 the fourth
# roots of the fourth (src/Contrl/Output.f90:330-341).
FORT_COLS = ((18, 0, 5), (24, 5, 6), (25, 11, 6), (26, 17, 6), (27, 23, 6), (29, 29, 6),
             (30, 35, 6))


def fort_row(d, k, prefix=""):
    """The 41-value diagnostic row k (0 = before the first step) assembled
    from a fixture's fort.18/24-27/29/30 arrays."""
    out = np.empty(41)
    for unit, lo, n in FORT_COLS:
        out[lo:lo + n] = d[f"{prefix}fort{unit}"][k, 1:1 + n]
    return out


def has_full_fort(d, prefix=""):
    return all(f"{prefix}fort{u}" in d for u, _, _ in FORT_COLS)


def full_row_err(got, ref):
    """Worst error over all 41 values of a diagnostic row.

    fort.24-26 as diag_row_err; every other value (fort.18's reference phase,
    gamma, energy, beta and r_max, the fort.27 maxima, the fort.29 cube roots
    and the fort.30 fourth roots) is normalised by its own magnitude, so a
    wrong exponent, sum index or scale factor shows at its full size.
    """
    g = np.asarray(got, dtype=np.float64)[:41]
    r = np.asarray(ref, dtype=np.float64)[:41]
    worst = diag_row_err(g, r)
    idx = np.r_[0:5, 23:41]
    scale = np.abs(r[idx])
    scale[scale == 0] = 1.0
    return max(worst, float(np.max(np.abs(g[idx] - r[idx]) / scale)))


def order_row_err(got, ref):
    """full_row_err for two rows whose sums differ only in the order of their
    terms.  A near-zero third moment's cube root (fort.29) magnifies a
    rounding-level change of its sum without bound, so those six values are
    compared as third moments, cubed, on the scale of the cube of the same
    coordinate's fourth-moment root (fort.30); every other value as in
    full_row_err."""
    g = np.asarray(got, dtype=np.float64)[:41].copy()
    r = np.asarray(ref, dtype=np.float64)[:41].copy()
    s3 = np.abs(r[35:41]) ** 3
    s3[s3 == 0] = 1.0
    third = float(np.max(np.abs(g[29:35] ** 3 - r[29:35] ** 3) / s3))
    g[29:35] = r[29:35]
    return max(full_row_err(g, r), third)


def fixture_row_err(got, d, k, prefix=""):
    """Error of a 41-value row against row k of a fixture: every column when
    the fixture holds all the diagnostic files, else fort.18 and 24-26."""
    got = np.asarray(got, dtype=np.float64)
    if has_full_fort(d, prefix):
        return full_row_err(got, fort_row(d, k, prefix))
    ref = np.concatenate([d[f"{prefix}fort{u}"][k, 1:] for u in (24, 25, 26)])
    r18 = d[f"{prefix}fort18"][k, 1:6]
    s18 = np.abs(r18)
    s18[s18 == 0] = 1.0
    return max(diag_row_err(got, ref, ref_is_fort=True),
               float(np.max(np.abs(got[:5] - r18) / s18)))


def deck_files_err(outdir, d, prefix=""):
    """Worst error of the diagnostic files a deck run wrote into outdir
    against a fixture's, every row and column: the z column absolutely, the
    41 values as fixture_row_err; fort.28 and fort.32 (where the fixture has
    them) must match exactly."""
    units = [u for u, _, _ in FORT_COLS if f"{prefix}fort{u}" in d]
    f = {u: np.loadtxt(os.path.join(outdir, f"fort.{u}"), ndmin=2) for u in units}
    nrows = d[f"{prefix}fort24"].shape[0]
    for u in units:
        assert f[u].shape == d[f"{prefix}fort{u}"].shape, (u, f[u].shape)
        assert np.max(np.abs(f[u][:, 0] - d[f"{prefix}fort{u}"][:, 0])) < 1e-12, u
    worst = 0.0
    for k in range(nrows):
        got = np.zeros(41)
        for u, lo, n in FORT_COLS:
            if u in f:
                got[lo:lo + n] = f[u][k, 1:1 + n]
        worst = max(worst, fixture_row_err(got, d, k, prefix))
    for u in (28, 32):
        if f"{prefix}fort{u}" in d:
            mine = np.loadtxt(os.path.join(outdir, f"fort.{u}"), ndmin=2)
            assert np.array_equal(mine[:, 1:], d[f"{prefix}fort{u}"][:, 1:]), u
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
