"""Pins of the dynamic Smagorinsky oracle and of sgdm's closed forms.

The reference has no dynamic model (solver/sgdm_incomp.f90:13-19 is its only
SGS model), so orc_dynsmag (oracle/emx_oracle.c) is pinned here by what the
mathematics fixes:

  * uniform flow and rigid rotation have S = 0, so M = 0 and nu_t = 0;
  * on a linear field u = A x of a grid with spacings h_a the box filter
    (1/4, 1/2, 1/4) adds h_a^2/2 to x_a^2 and leaves products of distinct
    coordinates alone, so L = A H^2 A^T / 2 (H = diag(h_a)); S is uniform,
    M = 2 delta^2 (1 - 4) |S| S, and nu_t = C delta^2 |S| =
    -(A H^2 A^T : S) / (12 S:S) at points two cells or more from the walls,
    whatever delta (h^2/12 for u = (x, y, -2z) on a uniform grid);
  * on small random fields, the model's definition written with full 3x3
    tensors and a 27-point filter (edge-clamped) instead of the oracle's six
    independent pairs and separable passes.

Each closed form is shown to fail when a constant of the model is planted
wrong in the oracle (filter ratio^2 2 instead of 4, filter weights 1/3
instead of 1/4-1/2-1/4, off-diagonal weight 1 instead of 2).

sgdm (viscosidad_t, solver/sgdm_incomp.f90:117-395) is pinned by the
flang-built fixtures; the closed forms below are its limits: uniform flow
gives F2 = 0; a rigid rotation has uniform vorticity, so SELECTIVO's angle
is 0 and amut = 0; a pure strain u = A x (A symmetric) has no vorticity
(the angle test compares NaN and keeps amut) and, on a uniform grid,
F2 = h^2 |A|_F^2 / 3, so amut = 0.1781 sqrt(F2) h.
"""
import numpy as np
import pytest

from conftest import normwise_rel

W = np.array([0.25, 0.5, 0.25])


def linear_case(n, spacing, A):
    """u = A x on an (n, n, n) grid with per-axis spacings; physical
    gradient g[3a+b] = A[a, b] everywhere."""
    hx, hy, hz = spacing
    i = np.arange(n, dtype=np.float64)
    Z, Y, X = np.meshgrid(i * hz, i * hy, i * hx, indexing="ij")
    xs = (X, Y, Z)
    u = np.stack([sum(A[a, b] * xs[b] for b in range(3)) for a in range(3)])
    g = np.empty((9, n, n, n))
    for a in range(3):
        for b in range(3):
            g[3 * a + b] = A[a, b]
    return np.ascontiguousarray(u), g


def closed_form(A, spacing):
    H2 = np.diag(np.asarray(spacing, dtype=np.float64) ** 2)
    S = 0.5 * (A + A.T)
    v = -np.sum((A @ H2 @ A.T) * S) / (12.0 * np.sum(S * S))
    return max(v, 0.0)


def interior(a, m=2):
    return a[m:-m, m:-m, m:-m]


def filt27(f):
    """27-point box filter, weights W x W x W, indices clamped at the walls."""
    p = np.pad(f, 1, mode="edge")
    nz, ny, nx = f.shape
    out = np.zeros_like(f)
    for c in range(3):
        for b in range(3):
            for a in range(3):
                out += W[c] * W[b] * W[a] * p[c:c + nz, b:b + ny, a:a + nx]
    return out


def dynsmag_tensor(u, g, delta):
    """The model's definition with full tensors (no symmetry shortcuts)."""
    G = g.reshape(3, 3, *g.shape[1:])           # G[a, b] = du_a/dx_b
    S = 0.5 * (G + G.transpose(1, 0, 2, 3, 4))
    aS = np.sqrt(2.0 * np.einsum("ij...,ij...->...", S, S))
    hS = np.stack([[filt27(S[i, j]) for j in range(3)] for i in range(3)])
    ahS = np.sqrt(2.0 * np.einsum("ij...,ij...->...", hS, hS))
    hSS = np.stack([[filt27(aS * S[i, j]) for j in range(3)] for i in range(3)])
    hu = np.stack([filt27(u[i]) for i in range(3)])
    L = np.stack([[filt27(u[i] * u[j]) - hu[i] * hu[j] for j in range(3)] for i in range(3)])
    M = 2.0 * delta ** 2 * (hSS - 4.0 * ahS * hS)
    LM = filt27(np.einsum("ij...,ij...->...", L, M))
    MM = filt27(np.einsum("ij...,ij...->...", M, M))
    C = np.where(MM > 0.0, LM / np.where(MM > 0.0, MM, 1.0), 0.0)
    return np.maximum(C * delta ** 2 * aS, 0.0)


def test_uniform_flow_and_rigid_rotation_give_zero(oracle):
    n = 10
    rng = np.random.default_rng(11)
    delta = rng.uniform(0.05, 0.2, (n, n, n))
    u, g = linear_case(n, (0.1, 0.1, 0.1), np.zeros((3, 3)))
    u += np.array([0.3, -1.2, 0.7])[:, None, None, None]
    assert np.array_equal(oracle.dynsmag(u, g, delta), np.zeros((n, n, n)))
    Om = np.array([[0.0, -0.4, 1.1], [0.4, 0.0, -0.7], [-1.1, 0.7, 0.0]])
    u, g = linear_case(n, (0.1, 0.13, 0.07), Om)
    # L = A H^2 A^T / 2 is not zero here: the model must still give zero
    assert np.abs(oracle.dynsmag(u, g, delta)).max() == 0.0


@pytest.mark.parametrize("delta", [0.01, 0.3])
def test_axisymmetric_strain_gives_h2_over_12(oracle, delta):
    """u = (x, y, -2z), spacing h: nu_t = h^2/12 whatever the filter width."""
    n, h = 9, 0.125
    A = np.diag([1.0, 1.0, -2.0])
    u, g = linear_case(n, (h, h, h), A)
    nut = oracle.dynsmag(u, g, np.full((n, n, n), delta))
    assert closed_form(A, (h, h, h)) == pytest.approx(h * h / 12.0, rel=1e-15)
    assert np.abs(interior(nut) / (h * h / 12.0) - 1.0).max() < 1e-12


def random_linear_fields(count, seed=5):
    """Non-symmetric A with off-diagonal terms and nu_t clearly > 0."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        A = rng.standard_normal((3, 3))
        A -= np.trace(A) / 3.0 * np.eye(3)
        sp = tuple(rng.uniform(0.05, 0.2, 3))
        if closed_form(A, sp) > 0.02 * float(np.mean(np.square(sp))):
            out.append((A, sp))
    return out


@pytest.mark.parametrize("case", range(4))
def test_linear_fields_anisotropic_spacing(oracle, case):
    """Any linear field on per-axis spacings: nu_t = -(A H^2 A^T : S)/(12 S:S)."""
    A, sp = random_linear_fields(4)[case]
    n = 8
    u, g = linear_case(n, sp, A)
    nut = oracle.dynsmag(u, g, np.full((n, n, n), 0.07))
    ex = closed_form(A, sp)
    assert np.abs(interior(nut) / ex - 1.0).max() < 1e-11


def test_definition_in_tensor_form_on_random_fields(oracle):
    """Random u, gradient and width on a small grid, walls included."""
    rng = np.random.default_rng(17)
    nz, ny, nx = 7, 8, 9
    u = rng.standard_normal((3, nz, ny, nx))
    g = rng.standard_normal((9, nz, ny, nx))
    delta = rng.uniform(0.05, 0.2, (nz, ny, nx))
    nut = oracle.dynsmag(u, g, delta)
    ref = dynsmag_tensor(u, g, delta)
    assert (ref > 0).mean() > 0.2 and (ref == 0).mean() > 0.05  # both branches
    assert normwise_rel(nut, ref) < 1e-11


def test_planted_mistakes_fail(oracle):
    """The checks above have teeth: each wrong constant fails at least one."""
    n, h = 9, 0.125
    A = np.diag([1.0, 1.0, -2.0])
    u, g = linear_case(n, (h, h, h), A)
    delta = np.full((n, n, n), 0.1)
    good = h * h / 12.0
    # test-filter ratio^2 of 2: h^2/4
    bad = interior(oracle.dynsmag(u, g, delta, ratio2=2.0))
    assert np.abs(bad / (h * h / 4.0) - 1.0).max() < 1e-12
    # filter weights 1/3, 1/3, 1/3: L grows by 4/3, nu_t = h^2/9
    bad = interior(oracle.dynsmag(u, g, delta, ws=1.0 / 3.0))
    assert np.abs(bad / good - 1.0).min() > 0.3
    # off-diagonal weight 1 in the contractions: the off-diagonal fields
    A2, sp = random_linear_fields(1)[0]
    u2, g2 = linear_case(8, sp, A2)
    d2 = np.full((8, 8, 8), 0.07)
    bad = interior(oracle.dynsmag(u2, g2, d2, offw=1.0))
    assert np.abs(bad / closed_form(A2, sp) - 1.0).min() > 1e-3
    # and every planted constant fails the tensor-form definition
    rng = np.random.default_rng(17)
    u3 = rng.standard_normal((3, 7, 8, 9))
    g3 = rng.standard_normal((9, 7, 8, 9))
    d3 = rng.uniform(0.05, 0.2, (7, 8, 9))
    ref = dynsmag_tensor(u3, g3, d3)
    for kw in ({"ratio2": 2.0}, {"ws": 1.0 / 3.0}, {"offw": 1.0}):
        assert normwise_rel(oracle.dynsmag(u3, g3, d3, **kw), ref) > 1e-3, kw


# ---- sgdm limits -------------------------------------------------------------

def sgdm_linear(oracle, A, n=10, L=1.0):
    """sgdm inputs for u = A x on a uniform n^3 grid of side L (metrics
    1/L, dvel the exact computational derivatives L*A)."""
    h = L / (n - 1)
    i = np.arange(n) * h
    Z, Y, X = np.meshgrid(i, i, i, indexing="ij")
    x = np.ascontiguousarray(np.stack([X, Y, Z]))
    u = np.ascontiguousarray(np.stack([A[a, 0] * X + A[a, 1] * Y + A[a, 2] * Z
                                       for a in range(3)]))
    dvel = np.empty((9, n, n, n))
    for b in range(3):
        for a in range(3):
            dvel[3 * b + a] = L * A[a, b]
    jbn = np.zeros((11, n, n, n))
    jbn[0] = jbn[4] = jbn[8] = 1.0 / L
    jbn[9] = L ** 3
    jbn[10] = 1.0 / L ** 3
    sgd, _, _ = oracle.inisgdm(x)
    return h, u, dvel, jbn, sgd


def test_sgdm_uniform_flow(oracle):
    h, u, dvel, jbn, sgd = sgdm_linear(oracle, np.zeros((3, 3)))
    u += 0.8
    amut, vis, _ = oracle.sgdm(np.ascontiguousarray(u), dvel, jbn, sgd, 1e-3)
    assert np.array_equal(amut, np.zeros_like(amut))
    assert np.array_equal(vis, np.full_like(vis, 1e-3))


def test_sgdm_rigid_rotation_switched_off(oracle):
    """Uniform vorticity: the angle to the neighbours' sum is zero, below
    20 degrees, so SELECTIVO zeroes amut although F2 > 0."""
    Om = np.array([[0.0, -0.5, 0.2], [0.5, 0.0, -0.9], [-0.2, 0.9, 0.0]])
    h, u, dvel, jbn, sgd = sgdm_linear(oracle, Om)
    amut, _, _ = oracle.sgdm(u, dvel, jbn, sgd, 1e-3)
    assert np.array_equal(amut, np.zeros_like(amut))


def test_sgdm_pure_strain_structure_function(oracle):
    """No vorticity: amut = 0.1781 * sqrt(h^2 |A|_F^2 / 3) * h at every
    point (one-sided differences of a linear field are exact, the filter
    factors are 1 on a uniform grid).  A wrong constant, a missing
    difference or a factor 1/3 exponent slip would break it."""
    A = np.array([[0.7, 0.3, -0.2], [0.3, -0.4, 0.5], [-0.2, 0.5, -0.3]])
    h, u, dvel, jbn, sgd = sgdm_linear(oracle, A, n=10, L=1.3)
    amut, vis, _ = oracle.sgdm(u, dvel, jbn, sgd, 1e-3)
    ex = 0.1781 * np.sqrt(h * h * np.sum(A * A) / 3.0) * h
    assert np.abs(amut / ex - 1.0).max() < 1e-12
    assert np.abs(vis - (1e-3 + amut)).max() < 1e-18
