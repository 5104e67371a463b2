"""Closed-form pins of the oracle, for what the Fortran fixtures leave open.

The C6C scheme, the level-set extrapolation and the WENO5/RK3 advection
exist only as CUDA in the reference (no nvcc here), so their restatements
are pinned by what the mathematics fixes; each check is chosen so that a
plausible slip (a wrong coefficient, sign, index or weight) fails it.
"""
import numpy as np
import pytest

from conftest import normwise_rel

C4, C6F, C6C = 0, 1, 2
DEGREE = {C4: 4, C6F: 3, C6C: 4}   # polynomial degree each scheme is exact to


def poly(x, deg, rng):
    c = rng.standard_normal(deg + 1)
    f = sum(c[k] * x ** k for k in range(deg + 1))
    df = sum(k * c[k] * x ** (k - 1) for k in range(1, deg + 1))
    return f, df


@pytest.mark.parametrize("scheme", [C4, C6F, C6C])
@pytest.mark.parametrize("n", [5, 6, 7, 9, 16, 33, 64])
def test_exact_on_polynomials_up_to_closure_order(oracle, scheme, n):
    """Every row of the scheme (interior Pade and both closures) is exact for
    polynomials up to its order, so the solved derivative is exact."""
    rng = np.random.default_rng(n + 10 * scheme)
    x = np.linspace(0.0, 1.0, n)
    f, df = poly(x, DEGREE[scheme], rng)
    got = oracle.deriv_line(scheme, f, float(n - 1))
    assert normwise_rel(got, df) < 1e-11


@pytest.mark.parametrize("scheme", [C4, C6F, C6C])
def test_not_exact_one_degree_higher(oracle, scheme):
    """...and not one degree higher: the test above has teeth."""
    n = 24
    x = np.linspace(0.0, 1.0, n)
    deg = DEGREE[scheme] + 1
    f, df = x ** deg, deg * x ** (deg - 1)
    got = oracle.deriv_line(scheme, f, float(n - 1))
    assert normwise_rel(got, df) > 1e-6


@pytest.mark.parametrize("scheme,order", [(C4, 4), (C6F, 6), (C6C, 6)])
def test_interior_convergence_order(oracle, scheme, order):
    """Away from the walls the error of d/dx sin falls as h^4 (C4, alpha=1/4,
    3/4) or h^6 (C6, alpha=1/3, 7/9, 1/36): this pins the interior
    coefficients beyond exactness on quartics."""
    errs = []
    for n in (65, 129):
        x = np.linspace(0.0, 1.0, n)
        k = 3.0
        f = np.sin(k * x + 0.3)
        got = oracle.deriv_line(scheme, f, float(n - 1))
        ex = k * np.cos(k * x + 0.3)
        mid = slice(n // 2 - 3, n // 2 + 4)
        errs.append(np.abs(got[mid] - ex[mid]).max())
    rate = np.log2(errs[0] / errs[1])
    assert order - 0.7 < rate < order + 1.5, (errs, rate)


@pytest.mark.parametrize("scheme", [C4, C6F, C6C])
def test_thomas_matches_dense_solve(oracle, scheme):
    """The derivative is A^{-1} B f: recover the operator D column by column
    and check A D is the banded matrix B the scheme writes (banded with the
    documented bandwidth, antisymmetric interior stencil)."""
    n = 12
    a, b, c = oracle.coeffs(scheme, n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    D = np.stack([oracle.deriv_line(scheme, np.eye(n)[j], 1.0) for j in range(n)], axis=1)
    B = A @ D
    # interior rows: antisymmetric stencil of half-width 2 (C6) or 1 (C4)
    hw = 1 if scheme == C4 else 2
    for l in range(3, n - 3):
        row = B[l]
        assert np.allclose(row[np.abs(np.arange(n) - l) > hw], 0.0, atol=1e-13)
        for d in range(1, hw + 1):
            assert row[l + d] == pytest.approx(-row[l - d], abs=1e-13)
        assert row[l] == pytest.approx(0.0, abs=1e-13)
    want = {C4: [0.75], C6F: [7 / 9, 1 / 36], C6C: [7 / 9, 1 / 36]}[scheme]
    for d, w in enumerate(want, start=1):
        assert B[n // 2, n // 2 + d] == pytest.approx(w, rel=1e-12)
    # the boundary closure rows: first row sums to zero (constants -> 0)
    assert abs(B[0].sum()) < 1e-12 and abs(B[-1].sum()) < 1e-12


def test_c6c_boundary_rows(oracle):
    """C6C closure (derivCompact.cu:109-140, 165-189): row 0 is
    f'_0 + 3 f'_1 = (-17/6, 3/2, 3/2, -1/6).f, row 1 is
    1/4 f'_0 + f'_1 + 1/4 f'_2 = 3/4 (f_2 - f_0); mirrored (negated) at the
    far end."""
    n = 10
    a, b, c = oracle.coeffs(C6C, n)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    D = np.stack([oracle.deriv_line(C6C, np.eye(n)[j], 1.0) for j in range(n)], axis=1)
    B = A @ D
    assert np.allclose(B[0, :4], [-17 / 6, 1.5, 1.5, -1 / 6], atol=1e-13)
    assert np.allclose(B[0, 4:], 0.0, atol=1e-13)
    assert np.allclose(B[1, :3], [-0.75, 0.0, 0.75], atol=1e-13)
    assert np.allclose(B[-1, -4:], [1 / 6, -1.5, -1.5, 17 / 6], atol=1e-13)
    assert np.allclose(B[-2, -3:], [-0.75, 0.0, 0.75], atol=1e-13)
    assert c[0] == 3.0 and a[-1] == 3.0 and a[1] == 0.25 and c[-2] == 0.25


# ---------------------------------------------------------------- level set

def padded_grid(n, shape_axis=0):
    """A padded uniform grid of n^3 points with spacing h and unit metric."""
    h = 1.0 / (n - 1)
    jbn = np.zeros((11, n, n, n))
    jbn[0] = jbn[4] = jbn[8] = 1.0
    return h, jbn


def test_weno5_exact_on_cubics(oracle):
    """Each 3-point ENO candidate is third-order, so for a cubic phi every
    convex combination -- whatever the nonlinear weights -- is exact:
    both one-sided derivatives equal phi'(x) away from the padded edges."""
    n = 20
    h, jbn = padded_grid(n)
    x = np.arange(n) * h
    for ax in range(3):
        shape = [1, 1, 1]
        shape[2 - ax] = n
        xx = x.reshape(shape)
        phi = np.ascontiguousarray(np.broadcast_to(0.3 + 2 * xx - 1.5 * xx ** 2 + 4 * xx ** 3,
                                                   (n, n, n)))
        dphi = np.broadcast_to(2 - 3 * xx + 12 * xx ** 2, (n, n, n))
        left, right = oracle.weno5_derivs(phi, jbn, (h, h, h))
        inner = [slice(4, n - 4)] * 3
        # left-biased uses D- at i-2..i+2: exact for i in 3..n-3
        assert normwise_rel(left[ax][tuple(inner)], dphi[tuple(inner)]) < 1e-11
        assert normwise_rel(right[ax][tuple(inner)], dphi[tuple(inner)]) < 1e-11


def test_weno5_fifth_order_on_smooth_data(oracle):
    """For smooth data the weights approach (0.1, 0.6, 0.3) and the
    combination is fifth order: a wrong linear weight (or swapped stencil)
    drops the rate to three."""
    errs = []
    for n in (40, 80):
        h, jbn = padded_grid(n)
        x = np.arange(n) * h
        phi = np.ascontiguousarray(np.broadcast_to(np.sin(2 * np.pi * x), (n, n, n)))
        left, right = oracle.weno5_derivs(phi, jbn, (h, h, h))
        ex = 2 * np.pi * np.cos(2 * np.pi * x)
        mid = slice(n // 4, 3 * n // 4)
        errs.append(max(np.abs(left[0, n // 2, n // 2, mid] - ex[mid]).max(),
                        np.abs(right[0, n // 2, n // 2, mid] - ex[mid]).max()))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 4.3, (errs, rate)


def test_weno5_left_right_mirror_symmetry(oracle):
    """The right-biased derivative of phi(x) equals minus the left-biased
    derivative of the mirrored field phi(-x) (interior points)."""
    n = 24
    h, jbn = padded_grid(n)
    rng = np.random.default_rng(5)
    line = np.cumsum(rng.standard_normal(n)) * 0.1
    phi = np.ascontiguousarray(np.broadcast_to(line, (n, n, n)))
    phim = np.ascontiguousarray(np.broadcast_to(line[::-1], (n, n, n)))
    left, right = oracle.weno5_derivs(phi, jbn, (h, h, h))
    leftm, _ = oracle.weno5_derivs(phim, jbn, (h, h, h))
    i = np.arange(4, n - 4)
    assert np.allclose(right[0, 5, 5, i], -leftm[0, 5, 5, n - 1 - i], rtol=1e-12, atol=1e-12)


def test_weno5_picks_the_smooth_stencil_at_a_jump(oracle):
    """ENO property: next to a step, the derivative is the smooth-side value
    (zero slope) to within the 1e-6 regularisation."""
    n = 24
    h, jbn = padded_grid(n)
    line = np.where(np.arange(n) < 12, 0.0, 1.0)
    phi = np.ascontiguousarray(np.broadcast_to(line, (n, n, n)))
    left, right = oracle.weno5_derivs(phi, jbn, (h, h, h))
    # at i=10 the left-biased stencil (D- at 8..12) contains the jump at 12:
    # the weights must favour the smooth candidates and give ~0
    assert abs(left[0, 5, 5, 10]) < 1e-3 / h
    assert abs(right[0, 5, 5, 13]) < 1e-3 / h


def test_weno5_rk3_translates_linear_phi_exactly(oracle):
    """Constant velocity, linear phi: WENO5 is exact, so one SSP-RK3 step is
    an exact translation phi - dt*u.grad(phi) in the interior (pins the
    upwind selection and the Shu-Osher coefficients 1, 3/4-1/4, 1/3-2/3)."""
    n = 30
    h, jbn = padded_grid(n)
    x = np.arange(n) * h
    g = (0.7, -1.3, 0.4)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    phi = np.ascontiguousarray(0.2 + g[0] * X + g[1] * Y + g[2] * Z)
    vel = np.zeros((3, n, n, n))
    vel[0], vel[1], vel[2] = 0.9, -0.5, 0.3
    dt = 0.3 * h
    out = oracle.ls_advect_weno5_padded(phi, vel, jbn, (h, h, h), dt)
    ex = phi - dt * (0.9 * g[0] - 0.5 * g[1] + 0.3 * g[2])
    s = slice(10, n - 10)
    assert normwise_rel(out[s, s, s], ex[s, s, s]) < 1e-13


def test_pad_mirror_is_the_even_reflection(oracle):
    """Ghost layer g-1-m equals interior layer m (cuGhostCellsMirror3d*)."""
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 7, 8))
    p = oracle.pad_mirror(a, 5)
    assert p.shape == (16, 17, 18)
    assert np.array_equal(p[5:11, 5:12, 5:13], a)
    for m in range(5):
        assert np.array_equal(p[5:11, 5:12, 4 - m], a[:, :, m])
        assert np.array_equal(p[5:11, 5:12, 13 + m], a[:, :, 7 - m])
        assert np.array_equal(p[5:11, 4 - m, 5:13], a[:, m, :])
        assert np.array_equal(p[4 - m, 5:12, 5:13], a[m, :, :])


def extrap_inputs(n):
    h = 1.0 / (n - 1)
    jbn = np.zeros((11, n, n, n))
    jbn[0] = jbn[4] = jbn[8] = 1.0
    return h, jbn


def test_extrapolation_keeps_constants(oracle):
    n = 16
    h, jbn = extrap_inputs(n)
    rng = np.random.default_rng(3)
    sob = rng.standard_normal((n, n, n))
    p = np.full((n, n, n), 3.25)
    out = oracle.ls_extrapolate(p, jbn, sob, (h, h, h), h)
    assert np.abs(out - 3.25).max() < 1e-14


def test_extrapolation_leaves_fluid_points(oracle):
    """rs vanishes where sobject <= 0 (so = [sobject > 0]): fluid points
    only see the RK combination, which returns them to within an ulp."""
    n = 14
    h, jbn = extrap_inputs(n)
    rng = np.random.default_rng(4)
    sob = rng.standard_normal((n, n, n))
    p = rng.standard_normal((n, n, n))
    out = oracle.ls_extrapolate(p, jbn, sob, (h, h, h), h)
    fl = sob <= 0
    assert np.abs(out[fl] - p[fl]).max() <= 4 * np.finfo(float).eps * np.abs(p).max()
    assert np.abs(out[~fl] - p[~fl]).max() > 1e-3


def test_extrapolation_is_upwind_transport_along_the_normal(oracle):
    """All solid (sobject = x), p linear in x: the normal is +x, the upwind
    (backward) difference of p is its slope b, so each RK3 stage moves p by
    -dt*b and 10 pseudo-steps give p - 10*dt*b with dt = deltamin/2
    (interface.cu:386-421; lsTransportEc.cu:170-171).  Points further than
    30 cells from the padded edges are untouched by the edge rules."""
    n = 80
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    jbn = np.zeros((11, 8, 8, n))
    jbn[0] = jbn[4] = jbn[8] = 1.0
    sob = np.ascontiguousarray(np.broadcast_to(0.5 + x, (8, 8, n)))
    b = 2.5
    p = np.ascontiguousarray(np.broadcast_to(1.0 + b * x, (8, 8, n)))
    dmin = 0.8 * h
    out = oracle.ls_extrapolate(p, jbn, sob, (h, 1.0 / 7, 1.0 / 7), dmin)
    mid = slice(35, n - 35)
    ex = p[:, :, mid] - 10 * (0.5 * dmin) * b
    assert normwise_rel(out[:, :, mid], ex) < 1e-13


def test_rk_update_is_ssp_rk3(oracle):
    """The three RUNGEK sub-step updates (solver/rungek_ico.f90:87-91,
    152-153, 216-219) applied to du/dt = lam*u give u0 * (1 + z + z^2/2 +
    z^3/6), z = lam*dt: the third-order Shu-Osher scheme.  A wrong
    coefficient in any sub-step breaks the cubic."""
    for z in (0.3, -0.7, 1.1):
        lam, dt = z / 0.05, 0.05
        u0 = np.array([1.0, 2.0, -0.5])
        zero = np.zeros_like(u0)
        u1 = oracle.rk_update(1, u0, None, lam * u0, zero, dt)
        u2 = oracle.rk_update(2, u0, u1, lam * u1, zero, dt)
        u3 = oracle.rk_update(3, u0, u2, lam * u2, zero, dt)
        ref = u0 * (1 + z + z * z / 2 + z ** 3 / 6)
        assert np.allclose(u3, ref, rtol=1e-14, atol=0)
    # rs and rsv enter as their sum
    a = oracle.rk_update(2, u0, u0 * 3, u0, u0 * 2, 0.1)
    b = oracle.rk_update(2, u0, u0 * 3, u0 * 3, zero, 0.1)
    assert np.allclose(a, b, rtol=1e-15)
