"""The drop-in, end to end: the reference's own Fortran call sites
(solver/deriv_vel_cuda.f90, solver/divergencia_cuda.f90, solver/advectLS.f90,
compiled unmodified from the reference sources by oracle/build_ref.sh) plus
the Fortran shim for flujos/viscosidad_t/viscosidad, linked against
libemix3d_hip.so, run one RUNGEK sub-step hot path on the GPU
(oracle/dropin_driver.f90).  Results are compared with the oracle at 1e-12
normwise (C6C, the scheme of the reference's CUDA path)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, normwise_rel
from emix3d_amd import fields

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "oracle", "_ref", "emx_dropin")
TOL = 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in program not built (oracle/_ref)")
def test_reference_call_sites_through_the_library(oracle, tmp_path):
    nx, ny, nz = 20, 18, 16
    f = fields.stage_fields(nx, ny, nz, 77, curvilinear=True, turbulent=True)
    z = {k: np.ascontiguousarray(v.numpy()) for k, v in f.items()}
    jbn = oracle.jacobeano(0, z["x"])
    sgd, dmins, _ = oracle.inisgdm(z["x"])
    sob = -fields.level_set_sphere_surface(f["x"], 5).numpy()
    dmin = float(dmins.min())
    c4 = 1.0 / fields.REYNOLDS
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fh)
        np.array([dmin, c4], dtype=np.float64).tofile(fh)
        for a in (z["u"], z["temp"], z["um"], z["pres"], jbn, sgd, sob):
            np.ascontiguousarray(a, dtype=np.float64).tofile(fh)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    N = nx * ny * nz
    sizes = [9, 3, 1, 1, 1, 5, 5, 1]
    parts = np.split(out, np.cumsum([s * N for s in sizes])[:-1])
    dvel, dtemp, amut, vis, vis5, rs, rsv, pres = (
        p.reshape((s, nz, ny, nx)) if s > 1 else p.reshape((nz, ny, nx))
        for p, s in zip(parts, sizes))
    # the Fortran passes 1/deltax and the C side inverts it again
    ih = tuple(1.0 / (1.0 / (n - 1)) for n in (nx, ny, nz))
    rdv, rdt = oracle.deriv_vel(2, z["u"], z["temp"], ih)
    assert normwise_rel(dvel, rdv) <= TOL
    assert normwise_rel(dtemp, rdt) <= TOL
    ra, rv, r5 = oracle.sgdm(z["u"], rdv, jbn, sgd, c4)
    assert normwise_rel(amut, ra) <= TOL
    assert normwise_rel(vis, rv) <= TOL
    assert normwise_rel(vis5, r5) <= TOL
    for neq in range(1, 6):
        flux = oracle.flujos(neq, z["um"], z["u"], z["pres"], rv, jbn, rdv)
        rrs, rrsv = oracle.divergencia(2, flux, jbn, ih)
        assert normwise_rel(rs[neq - 1], rrs) <= TOL, neq
        assert normwise_rel(rsv[neq - 1], rrsv) <= TOL, neq
    dx = tuple(1.0 / (n - 1) for n in (nx, ny, nz))
    rp = oracle.ls_extrapolate(z["pres"], jbn, np.ascontiguousarray(sob), dx, dmin)
    assert normwise_rel(pres, rp) <= TOL
