"""Adjoint kernels (transposed compact solve, divergence and flux adjoints)
through the C ABI.

  * parity with the oracle's adjoints (pinned in test_oracle_adjoint.py) on
    the same seeded inputs, 1e-12 normwise relative L-infinity;
  * the dot-product transpose test <w, D v> = <D^T w, v> to 1e-13 relative
    on the 128^3 stretched grid, for the derivative along every direction
    and for the composed operator (gradient + fluxes + divergence);
  * central finite differences (h = 1e-6) of the composed GPU operator on
    16 seeded random directions against the adjoint's directional
    derivative;
  * torch.autograd.gradcheck of the harness wrappers on a tiny grid.
"""
import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Operator, fields

pytestmark = pytest.mark.gpu
TOL = 1e-12
SID = {"c4": 0, "c6f": 1, "c6c": 2}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def report(name, err, tol=TOL):
    print(f"  {name}: {err:.2e}", flush=True)
    assert err <= tol, (name, err)


@pytest.mark.parametrize("scheme", ["c4", "c6f", "c6c"])
@pytest.mark.parametrize("shape", [(5, 6, 7), (13, 9, 33), (64, 48, 40), (100, 17, 129),
                                   (33, 257, 9), (512, 6, 5)])
def test_derivative_adj_vs_oracle(oracle, scheme, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape) + 1)
    g = rng.standard_normal((nz, ny, nx))
    op = Operator(nx, ny, nz, scheme)
    gd = dev(g)
    for d in range(3):
        ih = float(shape[d] - 1)
        got = host(op.derivative_adj(gd, d, ih))
        ref = oracle.deriv3d_adj(SID[scheme], d, g, ih)
        report(f"{scheme} {shape} D^T dir{d}", normwise_rel(got, ref))


def test_adjoint_rejects_long_lines():
    from emix3d_amd._lib import EmxError
    op = Operator(600, 5, 5, "c6c")
    g = torch.zeros((5, 5, 600), dtype=torch.float64, device="cuda")
    with pytest.raises(EmxError) as e:
        op.derivative_adj(g, 0)
    assert e.value.status == -3


def state(shape, seed, scheme, oracle, stretched=True):
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, seed, curvilinear=stretched, turbulent=False)
    z = {k: np.ascontiguousarray(v.numpy()) for k, v in f.items()}
    z["jbn"] = oracle.jacobeano(SID[scheme], z["x"])
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    z["dvel"], _ = oracle.deriv_vel(SID[scheme], z["u"], z["temp"], ih)
    rng = np.random.default_rng(seed + 1)
    z["vis"] = 1e-3 + 1e-3 * rng.random((nz, ny, nx))
    return z, ih


@pytest.mark.parametrize("scheme", ["c6f", "c6c"])
def test_divergence_and_flux_adj_vs_oracle(oracle, scheme):
    shape = (26, 21, 19)
    nx, ny, nz = shape
    z, ih = state(shape, 41, scheme, oracle)
    rng = np.random.default_rng(43)
    rb = rng.standard_normal((5, nz, ny, nx))
    rvb = rng.standard_normal((5, nz, ny, nx))
    op = Operator(*shape, scheme)
    fb = host(op.divergencia_adj(dev(rb[1]), dev(rvb[1]), dev(z["jbn"])))
    rfb = oracle.divergencia_adj(SID[scheme], np.ascontiguousarray(rb[1]),
                                 np.ascontiguousarray(rvb[1]), z["jbn"], ih)
    for q in range(6):
        report(f"flux_bar{q}", normwise_rel(fb[q], rfb[q]))
    bars = op.flujos_divergencia_all_adj(dev(rb), dev(rvb), dev(z["um"]), dev(z["u"]),
                                         dev(z["vis"]), dev(z["jbn"]), dev(z["dvel"]))
    ref = oracle.flux_div_all_adj(SID[scheme], rb, rvb, z["um"], z["u"], z["vis"], z["jbn"],
                                  z["dvel"], ih)
    for k in ("um_bar", "u_bar", "pres_bar", "vis_bar", "dvel_bar"):
        g, r = host(bars[k]), ref[k]
        if g.ndim == 4:
            for c in range(g.shape[0]):
                report(f"{k}[{c}]", normwise_rel(g[c], r[c]))
        else:
            report(k, normwise_rel(g, r))
    dvb = rng.standard_normal((9, nz, ny, nx))
    dtb = rng.standard_normal((3, nz, ny, nx))
    ub, tb = op.deriv_vel_adj(dev(dvb), dev(dtb))
    rub, rtb = oracle.deriv_vel_adj(SID[scheme], dvb, dtb, ih)
    report("grad_adj u_bar", normwise_rel(host(ub), rub))
    report("grad_adj temp_bar", normwise_rel(host(tb), rtb))


def gdot(a, b):
    return float(torch.sum(a * b).item())


def test_dot_product_128_stretched():
    """<w, D v> = <D^T w, v> at 128^3 for the derivative along every
    direction (metric-free), and for the divergence with the metrics of the
    tanh-stretched grid (beta = 2): <(rs_bar, rsv_bar), div(flux)> =
    <div^T(rs_bar, rsv_bar), flux>."""
    n = 128
    op = Operator(n, n, n, "c6c")
    gen = torch.Generator(device="cuda").manual_seed(20261206)
    for d in range(3):
        v = torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
        w = torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
        lhs = gdot(w, op.derivative(v, d))
        rhs = gdot(op.derivative_adj(w, d), v)
        report(f"dot dir{d}", abs(lhs - rhs) / max(abs(lhs), abs(rhs)), 1e-13)
    jbn = op.jacobeano(fields.grid_tanh(n, n, n, device="cuda"))
    assert float(jbn[0].max() / jbn[0].min()) > 2.0  # stretched
    flux = torch.randn((6, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rb = torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rvb = torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rs, rsv = op.divergencia(flux, jbn)
    lhs = gdot(rb, rs) + gdot(rvb, rsv)
    rhs = gdot(op.divergencia_adj(rb, rvb, jbn), flux)
    report("dot divergence (stretched metrics)", abs(lhs - rhs) / max(abs(lhs), abs(rhs)), 1e-13)


def composed(op, jbn, x):
    dvel, _ = op.deriv_vel(x["u"])
    return op.flujos_divergencia_all(x["um"], x["u"], x["pres"], x["vis"], jbn, dvel)


def composed_adj(op, jbn, x, rb, rvb):
    dvel, _ = op.deriv_vel(x["u"])
    b = op.flujos_divergencia_all_adj(rb, rvb, x["um"], x["u"], x["vis"], jbn, dvel)
    ub, _ = op.deriv_vel_adj(b["dvel_bar"])
    b["u_bar"] = b["u_bar"] + ub
    return b


def test_composed_operator_dot_fd_and_reference_128(oracle):
    """The flux divergence of (um, u, pres, vis) with dvel = grad u, on the
    128^3 stretched grid: dot-product test (the operator is quadratic, so
    (F(x+v) - F(x-v))/2 is its exact linearisation), 16 central finite
    differences with h = 1e-6, and the reference gradient of the CPU
    operator's transpose on a 40^3 cut of the same field family."""
    n = 128
    f = fields.stage_fields(n, n, n, 20261206, curvilinear=True, turbulent=False, device="cuda")
    op = Operator(n, n, n, "c6c")
    jbn = op.jacobeano(f["x"])
    gen = torch.Generator(device="cuda").manual_seed(4)
    x = {"um": f["um"], "u": f["u"], "pres": f["pres"],
         "vis": 1e-3 + 1e-3 * torch.rand((n, n, n), dtype=torch.float64, device="cuda",
                                         generator=gen)}
    rb = torch.randn((5, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rvb = torch.randn((5, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    bars = composed_adj(op, jbn, x, rb, rvb)
    names = ("um", "u", "pres", "vis")
    bar_of = {"um": "um_bar", "u": "u_bar", "pres": "pres_bar", "vis": "vis_bar"}
    worst_fd = 0.0
    for it in range(17):
        v = {k: torch.randn(x[k].shape, dtype=torch.float64, device="cuda", generator=gen)
             for k in names}
        h = 1.0 if it == 0 else 1e-6
        xp = {k: x[k] + h * v[k] for k in names}
        xm = {k: x[k] - h * v[k] for k in names}
        rp, rvp = composed(op, jbn, xp)
        rm, rvm = composed(op, jbn, xm)
        lhs = (gdot(rb, rp - rm) + gdot(rvb, rvp - rvm)) / (2.0 * h)
        rhs = sum(gdot(bars[bar_of[k]], v[k]) for k in names)
        err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        if it == 0:
            report("dot (exact linearisation)", err, 1e-13)
        else:
            worst_fd = max(worst_fd, err)
    # rounding of F / h bounds the finite-difference error: ~1e-16 * |F| / (h |L v|)
    report("16 central differences, h=1e-6 (worst)", worst_fd, 1e-7)
    # reference gradient: the oracle's transpose on a smaller grid of the family
    m = 40
    z, ih = state((m, m, m), 20261207, "c6c", oracle)
    rng = np.random.default_rng(5)
    rbs = rng.standard_normal((5, m, m, m))
    rvbs = rng.standard_normal((5, m, m, m))
    ref = oracle.flux_div_all_adj(2, rbs, rvbs, z["um"], z["u"], z["vis"], z["jbn"], z["dvel"], ih)
    rub, _ = oracle.deriv_vel_adj(2, ref["dvel_bar"], None, ih)
    ref["u_bar"] = ref["u_bar"] + rub
    ops = Operator(m, m, m, "c6c")
    xs = {k: dev(z[k]) for k in ("um", "u", "pres", "vis")}
    got = composed_adj(ops, dev(z["jbn"]), xs, dev(rbs), dev(rvbs))
    for k in ("um_bar", "u_bar", "pres_bar", "vis_bar"):
        report(f"reference {k}", normwise_rel(host(got[k]), ref[k]))


def test_autograd_gradcheck_tiny():
    from autograd_ops import Derivative, FluxDivOfState
    shape = (5, 6, 5)
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 3, curvilinear=True, turbulent=False, device="cuda")
    op = Operator(nx, ny, nz, "c6c")
    jbn = op.jacobeano(f["x"])
    um = f["um"].clone().requires_grad_(True)
    u = f["u"].clone().requires_grad_(True)
    pres = f["pres"].clone().requires_grad_(True)
    vis = (1e-2 + 0 * f["pres"]).clone().requires_grad_(True)
    # the operator is quadratic, so central differences are exact but for
    # rounding, ~1e-16 |F| / eps with |F| up to ~1e4 here
    assert torch.autograd.gradcheck(lambda a, b, c, d: FluxDivOfState.apply(op, jbn, a, b, c, d),
                                    (um, u, pres, vis), eps=1e-3, atol=1e-6, rtol=1e-6)
    g = torch.randn((nz, ny, nx), dtype=torch.float64, device="cuda").requires_grad_(True)
    for d in range(3):
        assert torch.autograd.gradcheck(lambda t: Derivative.apply(op, t, d), (g,), eps=1e-3,
                                        atol=1e-9, rtol=1e-7)
