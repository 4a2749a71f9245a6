"""The Fortran 90 host (impact-z_amd/fortran) against the reference.

izh_lattice calls the C ABI through the iso_c_binding module in the order of
run_AccSimulator's segment loop (or through izh_run); its per-step moment sums
must reproduce the reference's fort.24/25/26 rows to 1e-10 relative."""
import os
import subprocess

import numpy as np
import pytest

from conftest import fixture_beam, fixture_lattice, fixture_row_err, load_golden
from test_gpu_parity import TOL, record

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "impact-z_amd", "fortran", "izh_lattice")


def _namelist(path, d, fused, out):
    bm = fixture_beam(d)
    lat = fixture_lattice(d)
    nx, ny, nz = (int(v) for v in d["grid"])
    rows = [f"  nx={nx}, ny={ny}, nz={nz}, precision=0, device=0,",
            f"  current={bm.current!r}, kinetic={bm.kinetic!r}, mass={bm.mass!r},",
            f"  charge={bm.charge!r}, freq={bm.freq!r}, phase0={bm.phase0!r},",
            f"  nelem={len(lat)}, fused={'.true.' if fused else '.false.'},",
            f"  flagmap={int(d['flagmap']) if 'flagmap' in d else 1},",
            f"  particles='{path}/particle.in', output='{out}',"]
    for k, e in enumerate(lat, start=1):
        vals = ", ".join(f"val({j},{k})={v!r}" for j, v in enumerate(e.values, start=1))
        rows.append(f"  elen({k})={e.length!r}, nseg({k})={e.nseg}, mapstp({k})={e.mapstp},"
                    f" bitype({k})={e.type}, {vals}, radius({k})={e.radius!r},")
    with open(os.path.join(path, "run.nml"), "w") as f:
        f.write("&izh_run_nml\n" + "\n".join(rows) + "\n/\n")


@pytest.mark.parametrize("name,fused", [("fodo_traj", False), ("fodo_traj", True),
                                        ("cf_sol", False), ("cf_sol", True),
                                        ("nl_cf_sol", False), ("rf", False), ("rf", True)])
def test_fortran_host_matches_reference(tmp_path, name, fused):
    from impactz import hip
    from impactz.beam import rfdata_name, write_particle_in, write_rfdata
    from impactz.sim import diag_row
    assert os.access(EXE, os.X_OK), "Fortran host not built (impactz._build.build)"
    d = load_golden(name)
    write_particle_in(str(tmp_path / "particle.in"), d["pts_in"])
    for e in fixture_lattice(d):
        if e.field is not None:
            write_rfdata(str(tmp_path / rfdata_name(int(e.v4 + 0.1))), e.field)
    out = str(tmp_path / "moments.txt")
    _namelist(str(tmp_path), d, fused, out)
    print(f"fortran host {name} fused={fused}", flush=True)
    subprocess.run([EXE, str(tmp_path / "run.nml")], cwd=str(tmp_path), check=True,
                   timeout=120)
    rec = np.loadtxt(out, ndmin=2)
    nsteps = int(d["nsteps"])
    assert rec.shape == (nsteps, 49)
    bm = fixture_beam(d)
    c = hip.physconst(bm.freq)
    worst = 0.0
    for k in range(nsteps):
        refpt = np.zeros(6)
        refpt[5], refpt[4] = -rec[k, 47], rec[k, 48]
        assert rec[k, 47] == pytest.approx(d["fort18"][k + 1, 2], rel=1e-13)
        row = diag_row(rec[k, 3:40], rec[k, 40:46], rec[k, 46], int(rec[k, 2]), refpt,
                       bm.mass, c.scxl, c.pi)
        worst = max(worst, fixture_row_err(row, d, k + 1))
    assert abs(rec[-1, 1] - d["fort24"][nsteps, 0]) < 1e-12
    record(test="fortran_host", case=name, fused=fused, steps=nsteps, moments=worst)
    assert worst < TOL
