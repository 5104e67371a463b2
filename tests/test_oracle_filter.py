"""The oracle's compact filter (INIFILTRO arrays, FILTRONPER, FILTRADO)
against the reference Fortran's golden vectors and closed forms.

tests/golden/ref_filter.npz holds FILTRADO(itf) of the flang-built reference
(src/MSVSingle/herramientas_incomp2.f90 + filtro_incomp.f90) for all six
direction orders, cfilt 0.49 and 0.45, and a 9x8x7 grid whose lines are all
boundary rows (tests/golden/make_golden.py).  Closed forms: the interior
right-hand-side weights sum to 1 + 2 cfilt and cancel the odd-even mode, so
the non-periodic filter keeps constants exactly; the Thomas recurrence equals
a dense solve of the stated matrices.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_filter.npz")
CASES = [("g12x11x11", (1, 2, 3, 4, 5, 6)), ("g21x12x13", (1,)), ("g9x8x7", (1, 4))]


def jbn_of(z, name):
    jac = z[f"{name}_jac"]
    jbn = np.zeros((11,) + jac.shape[1:])
    jbn[9:11] = jac
    return jbn


@pytest.mark.parametrize("name,itfs", CASES)
def test_filtrado_matches_reference(oracle, name, itfs):
    z = np.load(GOLD)
    um1, cfilt, keep = z[f"{name}_um1"], float(z[f"{name}_cfilt"]), list(z[f"{name}_keep"])
    jbn = jbn_of(z, name)
    for itf in itfs:
        f, u, pres, temp = oracle.filtrado(itf, um1, jbn, cfilt, 0.3)
        assert np.array_equal(f[keep], z[f"{name}_itf{itf}_um1"]), (name, itf)
        if itf == itfs[0]:
            assert np.array_equal(u, z[f"{name}_itf{itf}_u"])
            assert np.array_equal(pres, z[f"{name}_itf{itf}_pres"])
            assert np.array_equal(temp, z[f"{name}_itf{itf}_temp"])
    # the direction orders give different results (the filter passes do not
    # commute near the boundary rows)
    if len(itfs) > 1:
        a = oracle.filtrado(1, um1, jbn, cfilt, 0.3)[0]
        b = oracle.filtrado(6, um1, jbn, cfilt, 0.3)[0]
        assert not np.array_equal(a, b)


def test_filter_rhs_weights():
    for cf in (0.0, 0.3, 0.45, 0.49):
        w = np.array([(193 + 126 * cf) / 256, (105 + 302 * cf) / 256, (-15 + 30 * cf) / 64,
                      (45 - 90 * cf) / 512, (-5 + 10 * cf) / 256, (1 - 2 * cf) / 512])
        # 0.5*(2 w0 + 2 w1 + ...) on a constant, and on (-1)^l
        assert abs(w.sum() - (1 + 2 * cf)) < 1e-15
        assert abs(w[0] - w[1] + w[2] - w[3] + w[4] - w[5]) < 1e-15


@pytest.mark.parametrize("n", [5, 9, 11, 16, 40])
def test_filtronper_constants_and_dense(oracle, n):
    cf = 0.47
    const = np.full(n, 2.5)
    np.testing.assert_allclose(oracle.filtronper(const, 0, cf), const, rtol=2e-15, atol=0)
    rng = np.random.default_rng(n)
    f = rng.standard_normal(n)
    if n < 11:
        # every row is a boundary row: identity matrix and identity RHS
        assert np.array_equal(oracle.filtronper(f, 0, cf), f)
    for periodic in (0, 1):
        a, b, c = oracle.filter_coeffs(periodic, n, cf)
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        B = np.eye(n)
        ws = [(193 + 126 * cf) / 256, (105 + 302 * cf) / 256, (-15 + 30 * cf) / 64,
              (45 - 90 * cf) / 512, (-5 + 10 * cf) / 256, (1 - 2 * cf) / 512]
        for l in range(5, n - 5):
            B[l] = 0.0
            B[l, l] = ws[0]
            for d in range(1, 6):
                B[l, l + d] += 0.5 * ws[d]
                B[l, l - d] += 0.5 * ws[d]
        ref = np.linalg.solve(A, B @ f)
        got = oracle.filtronper(f, periodic, cf)
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), (n, periodic)
