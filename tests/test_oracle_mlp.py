"""Pins of the oracle's closure network and features (no reference exists).

  * orc_mlp_grad against central differences of orc_mlp_forward (FP64),
    parameter by parameter on a sample of every block (weights and biases
    of all four layers): a dropped chain-rule factor or a transposed weight
    fails;
  * closed forms of orc_mlp_forward: with the hidden weights zero the output
    is out_scale * softplus(b4 + sum_k w4_k silu(b3_k)); softplus and SiLU
    at known points;
  * orc_mlp_features on a linear velocity field u = A x over a stretched
    grid: the compact derivative is exact for linear data, so feature 3a+b
    equals A[a][b] everywhere (fixes the index order), feature 9 the width.
"""
import math

import numpy as np
import pytest

from conftest import normwise_rel

NP = 9089
W1, B1, W2, B2, W3, B3, W4, B4 = 0, 640, 704, 4800, 4864, 8960, 9024, 9088


def rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(NP)).astype(np.float32)


def rand_norm(seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.standard_normal(10), 0.5 + rng.random(10), [0.7]]).astype(np.float32)


def test_mlp_grad_matches_central_differences(oracle):
    P = rand_params(1)
    norm = rand_norm(2)
    rng = np.random.default_rng(3)
    feat = rng.standard_normal((7, 10))
    dnu = rng.standard_normal(7)
    g = oracle.mlp_grad(P, norm, feat, dnu)
    picks = []
    for lo, hi in ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3), (B3, W4), (W4, B4), (B4, NP)):
        picks += list(rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False))
    for k in picks:
        h = 2.0 ** -12
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += np.float32(h)
        Pm[k] -= np.float32(h)
        fd = (oracle.mlp_forward(Pp, norm, feat) - oracle.mlp_forward(Pm, norm, feat)) / (2 * h)
        ref = float(np.dot(dnu, fd))
        assert abs(g[k] - ref) <= 1e-6 * max(1.0, abs(ref)), (k, g[k], ref)


def test_mlp_forward_closed_form(oracle):
    rng = np.random.default_rng(5)
    P = np.zeros(NP, np.float32)
    P[B3:W4] = rng.standard_normal(64)
    P[W4:B4] = rng.standard_normal(64)
    P[B4] = 0.25
    P[W1:B1] = rng.standard_normal(640)  # irrelevant once W2 = W3 = 0
    norm = rand_norm(6)
    feat = rng.standard_normal((4, 10))
    silu = lambda z: z / (1.0 + math.exp(-z))  # noqa: E731
    b3 = P[B3:W4].astype(np.float64)
    w4 = P[W4:B4].astype(np.float64)
    z = 0.25 + sum(w4[k] * silu(b3[k]) for k in range(64))
    ref = float(norm[20]) * math.log1p(math.exp(z))
    got = oracle.mlp_forward(P, norm, feat)
    assert np.allclose(got, ref, rtol=1e-14, atol=0)
    # softplus(0) = log 2 with every weight zero
    got0 = oracle.mlp_forward(np.zeros(NP, np.float32), norm, feat)
    assert np.allclose(got0, float(norm[20]) * math.log(2.0), rtol=1e-15)


@pytest.mark.parametrize("scheme", [0, 2])
def test_features_of_linear_field(oracle, scheme):
    from emix3d_amd import fields
    n = (12, 10, 9)
    x = fields.grid_tanh(*n).numpy().copy()
    A = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.9, 0.5, 1.1]])
    u = np.einsum("ab,bzyx->azyx", A, x).copy()
    jbn = oracle.jacobeano(scheme, x)
    ih = (n[0] - 1.0, n[1] - 1.0, n[2] - 1.0)
    dvel, _ = oracle.deriv_vel(scheme, u, u[0].copy(), ih)
    sgd, _, _ = oracle.inisgdm(x)
    delta = np.ascontiguousarray(sgd[0])
    f = oracle.mlp_features(dvel, jbn, delta)
    for a in range(3):
        for b in range(3):
            assert normwise_rel(f[:, 3 * a + b], np.full(f.shape[0], A[a, b])) < 1e-10, (a, b)
    assert np.array_equal(f[:, 9], delta.reshape(-1))
