"""Pins of the oracle's adjoints (the reference has none).

Each adjoint must be the transpose of the oracle's forward function, which
the golden fixtures and closed forms pin to the reference.  Checks:
  * the transposed dense operator matrix on short lines (built column by
    column from the forward derivative);
  * dot-product tests <w, L v> = <L^T w, v> for the derivative along every
    direction, the divergence, and the linearised five-equation flux
    divergence (alone and composed with the velocity gradient).  The flux
    divergence is quadratic in its inputs, so the central difference
    (F(x+v) - F(x-v))/2 is its exact linearisation L v (no truncation);
  * teeth: a transposed-but-wrong adjoint (the forward itself, or one flux
    term dropped) fails the same checks.
"""
import numpy as np
import pytest

from conftest import normwise_rel

C4, C6F, C6C = 0, 1, 2
TOL_DOT = 1e-13


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("scheme", [C4, C6F, C6C])
@pytest.mark.parametrize("n", [5, 6, 9, 17])
def test_deriv_adj_is_dense_transpose(oracle, scheme, n):
    ih = float(n - 1)
    D = np.stack([oracle.deriv_line(scheme, np.eye(n)[l], ih) for l in range(n)], axis=1)
    DT = np.stack([oracle.deriv3d_adj(scheme, 0, np.eye(n)[l].reshape(1, 1, n), ih).ravel()
                   for l in range(n)], axis=1)
    assert normwise_rel(DT, D.T) < 1e-14
    # teeth: the operator is not symmetric, so the forward is no adjoint
    assert normwise_rel(D, D.T) > 1e-3


@pytest.mark.parametrize("scheme", [C4, C6F, C6C])
def test_deriv3d_adj_dot_product(oracle, scheme):
    rng = np.random.default_rng(7 + scheme)
    shape = (9, 11, 13)  # (nz, ny, nx)
    for d in range(3):
        n = shape[2 - d]
        v = rng.standard_normal(shape)
        w = rng.standard_normal(shape)
        lhs = float(np.sum(w * oracle.deriv3d(scheme, d, v, float(n - 1))))
        rhs = float(np.sum(oracle.deriv3d_adj(scheme, d, w, float(n - 1)) * v))
        assert rel(lhs, rhs) < TOL_DOT, (d, lhs, rhs)


def stage_state(shape, seed, oracle, scheme=C6C):
    from emix3d_amd import fields
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, seed, curvilinear=True, turbulent=True)
    z = {k: np.ascontiguousarray(v.numpy()) for k, v in f.items()}
    z["jbn"] = oracle.jacobeano(scheme, z["x"])
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    z["dvel"], _ = oracle.deriv_vel(scheme, z["u"], z["temp"], ih)
    rng = np.random.default_rng(seed)
    z["vis"] = 1e-3 + 1e-3 * rng.random((nz, ny, nx))
    return z, ih


def flux_div_all(oracle, scheme, um, u, pres, vis, jbn, dvel, ih):
    rs, rsv = [], []
    for m in range(1, 6):
        fl = oracle.flujos(m, um, u, pres, vis, jbn, dvel)
        a, b = oracle.divergencia(scheme, fl, jbn, ih)
        rs.append(a)
        rsv.append(b)
    return np.stack(rs), np.stack(rsv)


@pytest.mark.parametrize("scheme", [C4, C6C])
def test_divergencia_adj_dot_product(oracle, scheme):
    shape = (14, 10, 12)
    z, ih = stage_state(shape, 3, oracle, scheme)
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    flux = rng.standard_normal((6, nz, ny, nx))
    rb, rvb = rng.standard_normal((2, nz, ny, nx))
    rs, rsv = oracle.divergencia(scheme, flux, z["jbn"], ih)
    lhs = float(np.sum(rb * rs) + np.sum(rvb * rsv))
    fb = oracle.divergencia_adj(scheme, np.ascontiguousarray(rb), np.ascontiguousarray(rvb),
                                z["jbn"], ih)
    rhs = float(np.sum(fb * flux))
    assert rel(lhs, rhs) < TOL_DOT


def linearised(oracle, scheme, z, ih, v, grad=False):
    """Exact L v of the (quadratic) five-equation flux divergence by the
    central difference with step 1; with grad, dvel = deriv_vel(u)."""
    def F(s):
        um = z["um"] + s * v["um"]
        u = z["u"] + s * v["u"]
        pres = z["pres"] + s * v["pres"]
        vis = z["vis"] + s * v["vis"]
        if grad:
            dvel, _ = oracle.deriv_vel(scheme, u, z["temp"], ih)
        else:
            dvel = z["dvel"] + s * v["dvel"]
        return flux_div_all(oracle, scheme, um, u, pres, vis, z["jbn"], dvel, ih)
    rp, rvp = F(1.0)
    rm, rvm = F(-1.0)
    return 0.5 * (rp - rm), 0.5 * (rvp - rvm)


def directions(shape, rng):
    nx, ny, nz = shape
    return {"um": rng.standard_normal((5, nz, ny, nx)), "u": rng.standard_normal((3, nz, ny, nx)),
            "pres": rng.standard_normal((nz, ny, nx)), "vis": rng.standard_normal((nz, ny, nx)),
            "dvel": rng.standard_normal((9, nz, ny, nx))}


def dot_bars(bars, v, grad=False):
    s = (np.sum(bars["um_bar"] * v["um"]) + np.sum(bars["u_bar"] * v["u"])
         + np.sum(bars["pres_bar"] * v["pres"]) + np.sum(bars["vis_bar"] * v["vis"]))
    if not grad:
        s += np.sum(bars["dvel_bar"] * v["dvel"])
    return float(s)


@pytest.mark.parametrize("scheme", [C6F, C6C])
def test_flux_div_all_adj_dot_product(oracle, scheme):
    shape = (12, 9, 10)
    z, ih = stage_state(shape, 5, oracle, scheme)
    rng = np.random.default_rng(13)
    v = directions(shape, rng)
    nx, ny, nz = shape
    rb = rng.standard_normal((5, nz, ny, nx))
    rvb = rng.standard_normal((5, nz, ny, nx))
    Lrs, Lrsv = linearised(oracle, scheme, z, ih, v)
    lhs = float(np.sum(rb * Lrs) + np.sum(rvb * Lrsv))
    bars = oracle.flux_div_all_adj(scheme, rb, rvb, z["um"], z["u"], z["vis"], z["jbn"],
                                   z["dvel"], ih)
    rhs = dot_bars(bars, v)
    assert rel(lhs, rhs) < TOL_DOT, (lhs, rhs)
    assert np.all(bars["um_bar"][0] == 0.0)


def test_flux_div_all_adj_composed_with_gradient(oracle):
    """Chain rule through dvel = deriv_vel(u): u_bar += deriv_vel_adj(dvel_bar)."""
    scheme = C6C
    shape = (10, 11, 9)
    z, ih = stage_state(shape, 9, oracle, scheme)
    rng = np.random.default_rng(17)
    v = directions(shape, rng)
    nx, ny, nz = shape
    rb = rng.standard_normal((5, nz, ny, nx))
    rvb = rng.standard_normal((5, nz, ny, nx))
    Lrs, Lrsv = linearised(oracle, scheme, z, ih, v, grad=True)
    lhs = float(np.sum(rb * Lrs) + np.sum(rvb * Lrsv))
    bars = oracle.flux_div_all_adj(scheme, rb, rvb, z["um"], z["u"], z["vis"], z["jbn"],
                                   z["dvel"], ih)
    ub, _ = oracle.deriv_vel_adj(scheme, bars["dvel_bar"], None, ih)
    bars["u_bar"] = bars["u_bar"] + ub
    rhs = dot_bars(bars, v, grad=True)
    assert rel(lhs, rhs) < TOL_DOT, (lhs, rhs)


def test_deriv_vel_adj_dot_product(oracle):
    scheme = C6F
    shape = (8, 12, 10)
    nx, ny, nz = shape
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    rng = np.random.default_rng(19)
    u = rng.standard_normal((3, nz, ny, nx))
    t = rng.standard_normal((nz, ny, nx))
    dvb = rng.standard_normal((9, nz, ny, nx))
    dtb = rng.standard_normal((3, nz, ny, nx))
    dvel, dtemp = oracle.deriv_vel(scheme, u, t, ih)
    lhs = float(np.sum(dvb * dvel) + np.sum(dtb * dtemp))
    ub, tb = oracle.deriv_vel_adj(scheme, dvb, dtb, ih)
    rhs = float(np.sum(ub * u) + np.sum(tb * t))
    assert rel(lhs, rhs) < TOL_DOT


def test_dot_product_has_teeth(oracle):
    """Dropping one term of the flux adjoint (the pressure adjoint) is caught."""
    scheme = C6C
    shape = (10, 9, 8)
    z, ih = stage_state(shape, 23, oracle, scheme)
    rng = np.random.default_rng(29)
    v = directions(shape, rng)
    nx, ny, nz = shape
    rb = rng.standard_normal((5, nz, ny, nx))
    rvb = rng.standard_normal((5, nz, ny, nx))
    Lrs, Lrsv = linearised(oracle, scheme, z, ih, v)
    lhs = float(np.sum(rb * Lrs) + np.sum(rvb * Lrsv))
    bars = oracle.flux_div_all_adj(scheme, rb, rvb, z["um"], z["u"], z["vis"], z["jbn"],
                                   z["dvel"], ih)
    bars["pres_bar"] = np.zeros_like(bars["pres_bar"])
    assert rel(lhs, dot_bars(bars, v)) > 1e-6
