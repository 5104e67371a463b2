"""The level-set oracle's metric coupling, pinned with per-axis metrics.

A uniform grid of physical side lengths (Lx, Ly, Lz) = (2.0, 0.5, 1.3) over
the unit computational cube has the inverse metrics xi_x = 1/Lx,
eta_y = 1/Ly, zeta_z = 1/Lz (jbn 1, 5, 9).  On fields that are linear in
the physical coordinates every difference the level-set kernels take is
exact, so:

  * WENO5 + SSP-RK3 advection (cuAdvectLsJB, src/LevelSet/lsTransportEc.cu:
    201-418, with PhiDev{Plus,Minus}ParameterJB, lsTools.cu:587-770) is the
    exact translation phi - dt u.grad(phi) with the physical slopes;
  * the pressure extrapolation (extrapolKernel and DevFirstOrder_LS,
    src/LevelSet/extrapol.cu:5-153, ten SSP-RK3 pseudo-steps of
    interface.cu:386-421) moves p by 10 dtau (n.grad p) with n the physical
    gradient of sobject (it is not normalised) and dtau = deltamin/2,
    upwinded per axis by the sign of n.

Planting a swapped metric (jbn 1 <-> jbn 5) or a dropped one (jbn 9 = 1)
in the inputs breaks both: the oracle must read each axis's own metric.
"""
import numpy as np
import pytest

from conftest import normwise_rel

LENGTHS = (2.0, 0.5, 1.3)


def grid(n):
    h = 1.0 / (n - 1)
    s = np.arange(n) * h
    Z, Y, X = np.meshgrid(LENGTHS[2] * s, LENGTHS[1] * s, LENGTHS[0] * s, indexing="ij")
    jbn = np.zeros((11, n, n, n))
    jbn[0], jbn[4], jbn[8] = 1.0 / LENGTHS[0], 1.0 / LENGTHS[1], 1.0 / LENGTHS[2]
    jbn[9] = LENGTHS[0] * LENGTHS[1] * LENGTHS[2]
    jbn[10] = 1.0 / jbn[9]
    return h, (X, Y, Z), jbn


def planted(jbn):
    swap = jbn.copy()
    swap[0], swap[4] = jbn[4], jbn[0]
    drop = jbn.copy()
    drop[8] = 1.0
    return {"swap 1<->5": swap, "drop 9": drop}


def weno_case(n=30):
    h, (X, Y, Z), jbn = grid(n)
    gphys = (0.7, -1.3, 0.4)
    phi = np.ascontiguousarray(0.2 + gphys[0] * X + gphys[1] * Y + gphys[2] * Z)
    uvel = (0.9, -0.5, 0.3)
    vel = np.empty((3, n, n, n))
    for c in range(3):
        vel[c] = uvel[c]
    dt = 0.3 * min(LENGTHS) * h
    ex = phi - dt * sum(uvel[c] * gphys[c] for c in range(3))
    return h, phi, vel, jbn, dt, ex


def test_weno_translation_with_physical_slopes(oracle):
    """On the padded array directly (no ghost construction involved)."""
    n = 30
    h, phi, vel, jbn, dt, ex = weno_case(n)
    out = oracle.ls_advect_weno5_padded(phi, vel, jbn, (h, h, h), dt)
    s = slice(10, n - 10)
    assert normwise_rel(out[s, s, s], ex[s, s, s]) < 1e-13
    for name, bad in planted(jbn).items():
        o = oracle.ls_advect_weno5_padded(phi, vel, bad, (h, h, h), dt)
        assert normwise_rel(o[s, s, s] - phi[s, s, s], ex[s, s, s] - phi[s, s, s]) > 1e-3, name


def test_weno_translation_through_the_mirror_padding(oracle):
    """The unpadded entry (5-cell mirror padding of phi, velocity and
    metrics) away from the faces, where the mirror's kink cannot reach
    within one RK3 step."""
    n = 30
    h, phi, vel, jbn, dt, ex = weno_case(n)
    out = oracle.ls_advect_weno5(phi, vel, jbn, (h, h, h), dt)
    s = slice(10, n - 10)
    assert normwise_rel(out[s, s, s], ex[s, s, s]) < 1e-13
    bad = planted(jbn)["swap 1<->5"]
    o = oracle.ls_advect_weno5(phi, vel, bad, (h, h, h), dt)
    assert normwise_rel(o[s, s, s] - phi[s, s, s], ex[s, s, s] - phi[s, s, s]) > 1e-3


@pytest.mark.parametrize("nvec", [(0.6, 0.8, 0.5), (0.6, -0.8, 0.5)])
def test_extrapolation_along_physical_normal(oracle, nvec):
    """All points solid, sobject = 5 + n.x and p = 1 + b.x linear in the
    physical coordinates.  The first-order gradient of sobject is exact;
    p's upwind difference along each axis (by the sign of that component
    of n) is exact; every RK3 stage moves p by -dtau n.b.  Within 30 cells
    downwind of a face the mirror's kink arrives, so only the middle of a
    72^3 grid is compared."""
    n = 72
    h, (X, Y, Z), jbn = grid(n)
    nv = 1.7 * np.asarray(nvec)
    b = np.array([0.9, -0.4, 1.6])
    sob = np.ascontiguousarray(5.0 + nv[0] * X + nv[1] * Y + nv[2] * Z)
    p = np.ascontiguousarray(1.0 + b[0] * X + b[1] * Y + b[2] * Z)
    dmin = 0.8 * min(LENGTHS) * h
    out = oracle.ls_extrapolate(p, jbn, sob, (h, h, h), dmin)
    mid = slice(33, 39)
    ex = p - 10 * (0.5 * dmin) * float(nv @ b)
    assert normwise_rel(out[mid, mid, mid] - p[mid, mid, mid],
                        ex[mid, mid, mid] - p[mid, mid, mid]) < 1e-11
    for name, bad in planted(jbn).items():
        o = oracle.ls_extrapolate(p, bad, sob, (h, h, h), dmin)
        assert normwise_rel(o[mid, mid, mid] - p[mid, mid, mid],
                            ex[mid, mid, mid] - p[mid, mid, mid]) > 1e-3, name
