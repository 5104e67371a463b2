"""The oracle's level-set reinitialisation (cuReinitLsJB) against what the
mathematics fixes.  The reference's CUDA cannot be built here, so:

  * the Z/Y/X ghost mirroring equals the direct mirror map of the padding
    (two independent codes);
  * a signed distance to a plane (|grad phi| = 1 exactly, reproduced exactly
    by WENO5 on a linear field) is a fixed point away from the box faces,
    on a uniform and on a stretched grid (the metrics enter the gradient);
  * on a distorted signed distance to a sphere, the sign of phi and the
    position of the zero level set are kept while | |grad phi| - 1 |
    drops near the interface, and the smoothed sign with dmin -> 0 makes no
    difference away from the interface.
A dropped metric, a wrong upwind branch or a wrong sign of S fails one of
them (checked below by planting such mistakes in copies of the inputs).
"""
import numpy as np

from emix3d_amd import fields


def uniform(n):
    x = fields.grid_uniform(*n).numpy()
    return x


def metrics(oracle, x):
    return oracle.jacobeano(2, np.ascontiguousarray(x))


def test_ghost_mirror_equals_padding(oracle):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((9, 11, 13))
    padded = oracle.pad_mirror(a)
    scrambled = np.array(padded)
    g = 5
    mask = np.ones_like(scrambled, dtype=bool)
    mask[g:-g, g:-g, g:-g] = False
    scrambled[mask] = rng.standard_normal(mask.sum())
    assert np.array_equal(oracle.ls_ghost_mirror(scrambled), padded)


def grad_norm(phi, x):
    """|grad phi| by second-order differences on the (possibly stretched)
    orthogonal grid."""
    gz = np.gradient(phi, x[2][:, 0, 0], axis=0)
    gy = np.gradient(phi, x[1][0, :, 0], axis=1)
    gx = np.gradient(phi, x[0][0, 0, :], axis=2)
    return np.sqrt(gx * gx + gy * gy + gz * gz)


def run(oracle, phi, x, dmin_scale=1.0):
    nz, ny, nx = phi.shape
    jbn = metrics(oracle, x)
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    h = min(np.diff(x[0][0, 0, :]).min(), np.diff(x[1][0, :, 0]).min(),
            np.diff(x[2][:, 0, 0]).min())
    dmin = np.full(phi.shape, dmin_scale * h)
    return oracle.ls_reinit(phi, jbn, dmin, dxyz, min(dxyz)), h


def plane(shape, stretched):
    x = (fields.grid_tanh(*shape, beta=1.2) if stretched else fields.grid_uniform(*shape)).numpy()
    return x, 0.6 * (x[0] - 0.45) + 0.8 * (x[2] - 0.55)


def test_plane_distance_is_a_fixed_point(oracle):
    """Uniform grid: WENO5 is exact on the linear field, so the plane
    distance is a fixed point to rounding.  The mirror ghost cells put a
    kink at every face, whose influence decays by about a factor 100 per
    two cells inward: 11 cells in, it is below rounding."""
    x, phi = plane((36, 34, 32), False)
    out, _ = run(oracle, phi, x)
    core = (slice(11, -11),) * 3
    assert np.abs(out - phi)[core].max() < 2e-12


def test_plane_distance_stretched_converges(oracle):
    """Stretched grid: the field is linear in x but not in the computational
    coordinate, so |grad phi| = 1 holds to the WENO5 and metric truncation
    error; the change in the core is small and falls with resolution (the
    metrics enter: without them the change is O(1))."""
    errs = []
    for n in (36, 56):
        x, phi = plane((n, n - 2, n - 4), True)
        out, _ = run(oracle, phi, x)
        core = (slice(11, -11),) * 3
        errs.append(np.abs(out - phi)[core].max())
    assert errs[0] < 1e-5 and errs[1] < 0.25 * errs[0], errs


def sphere(shape, distort):
    x = fields.grid_uniform(*shape).numpy()
    r = np.sqrt(((x - 0.5) ** 2).sum(0))
    d = r - 0.3
    return x, d, d * (1.0 + distort * d / 0.3) * (1.5 if distort else 1.0)


def test_distorted_sphere_is_repaired(oracle):
    shape = (40, 40, 40)
    x, d, phi = sphere(shape, 2.0)
    out, h = run(oracle, phi, x)
    band = np.abs(d) < 3 * h
    inner = (slice(3, -3),) * 3
    before = np.abs(grad_norm(phi, x) - 1.0)[inner][band[inner]].mean()
    after = np.abs(grad_norm(out, x) - 1.0)[inner][band[inner]].mean()
    assert after < 0.3 * before, (before, after)
    # sign kept away from the interface, interface held within a tenth of a cell
    far = np.abs(d) > 2 * h
    assert np.array_equal(np.sign(out[far]), np.sign(phi[far]))
    near = np.abs(d) < 0.5 * h
    assert np.abs(out[near] - d[near]).max() < 0.1 * h * 1.5 + np.abs(phi - d)[near].max()


def test_planted_mistakes_fail(oracle):
    """The fixed-point test has teeth: a metric off by 10% moves the plane
    distance, and a reversed sign of phi0 is not a fixed point either."""
    shape = (30, 28, 26)
    x = fields.grid_uniform(*shape).numpy()
    phi = 0.6 * (x[0] - 0.45) + 0.8 * (x[2] - 0.55)
    nz, ny, nx = phi.shape
    jbn = metrics(oracle, x)
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    dmin = np.full(phi.shape, 1.0 / 30)
    bad = np.array(jbn)
    bad[0] *= 1.1
    out = oracle.ls_reinit(phi, bad, dmin, dxyz, min(dxyz))
    core = (slice(9, -9), slice(9, -9), slice(9, -9))
    assert np.abs(out - phi)[core].max() > 1e-4
    # a field with |grad phi| = 2 moves toward the distance: also not fixed
    out2 = oracle.ls_reinit(2 * phi, jbn, dmin, dxyz, min(dxyz))
    assert np.abs(out2 - 2 * phi)[core].max() > 1e-4
