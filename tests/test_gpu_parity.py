"""HIP kernels (through the C ABI) against the CPU oracle on the same seeded
inputs.  The oracle is pinned to the reference Fortran by
test_oracle_golden.py and to closed forms by test_oracle_closed_forms.py.

Tolerance: 1e-12 normwise relative L-infinity per output (max|gpu-ref| /
max|ref|), the bound the operator is specified to meet; the observed errors
are printed and are typically 1e-16..1e-15.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Operator, fields

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCHEMES = ["c4", "c6f", "c6c"]
SID = {"c4": 0, "c6f": 1, "c6c": 2}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def report(name, err):
    print(f"  {name}: rel err {err:.2e}", flush=True)
    assert err <= TOL, (name, err)


def inputs(shape, seed, curv=True, turb=True):
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, seed, curvilinear=curv, turbulent=turb)
    return {k: np.ascontiguousarray(v.numpy()) for k, v in f.items()}


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", [(5, 6, 7), (13, 9, 33), (64, 64, 64), (100, 17, 129),
                                   (256, 20, 12), (33, 257, 9)])
def test_derivative_all_directions(oracle, scheme, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal((nz, ny, nx))
    op = Operator(nx, ny, nz, scheme)
    fd = dev(f)
    for d in range(3):
        ih = float(shape[d] - 1)
        got = host(op.derivative(fd, d, ih))
        ref = oracle.deriv3d(SID[scheme], d, f, ih)
        report(f"{scheme} {shape} dir{d}", normwise_rel(got, ref))


PLANS = [(0, [1, 2, 4, 8, 16], 64), (1, [8, 16, 32], 16), (2, [4, 8, 16], 32),
         (3, [4, 8], 64), (4, [4, 8], 128)]


@pytest.mark.parametrize("kind,chunks,pmax", PLANS)
def test_derivative_every_plan(oracle, kind, chunks, pmax):
    """Every line-solver shape the autotuner may pick gives the same result
    (sizes not multiples of the chunk or of the lines per wave)."""
    nx, ny, nz = 96, 81, 70
    rng = np.random.default_rng(3)
    f = rng.standard_normal((nz, ny, nx))
    op = Operator(nx, ny, nz, "c6c")
    fd = dev(f)
    for d in range(3):
        n = (nx, ny, nz)[d]
        ref = oracle.deriv3d(2, d, f, float(n - 1))
        for C in chunks:
            P = -(-n // C)
            if P > pmax:
                continue
            op.set_plan(d, kind, C)
            got = host(op.derivative(fd, d))
            report(f"kind{kind} C{C} dir{d}", normwise_rel(got, ref))


@pytest.mark.parametrize("scheme", SCHEMES)
def test_deriv_vel(oracle, scheme):
    shape = (24, 20, 18)
    z = inputs(shape, 11)
    op = Operator(*shape, scheme)
    ih = op.default_ih()
    dvel, dtemp = op.deriv_vel(dev(z["u"]), dev(z["temp"]))
    rdv, rdt = oracle.deriv_vel(SID[scheme], z["u"], z["temp"], ih)
    dvel, dtemp = host(dvel), host(dtemp)
    for c in range(9):
        report(f"dvel{c + 1}", normwise_rel(dvel[c], rdv[c]))
    for c in range(3):
        report(f"dtemp{c + 1}", normwise_rel(dtemp[c], rdt[c]))
    dvel2, _ = op.deriv_vel(dev(z["u"]))
    report("dvel without temp", normwise_rel(host(dvel2), rdv))


@pytest.mark.parametrize("scheme", ["c4", "c6c"])
def test_jacobeano_and_inisgdm(oracle, scheme):
    shape = (20, 16, 24)
    z = inputs(shape, 5)
    op = Operator(*shape, scheme)
    jbn = host(op.jacobeano(dev(z["x"])))
    rj = oracle.jacobeano(SID[scheme], z["x"])
    for c in range(11):
        report(f"jbn{c + 1}", normwise_rel(jbn[c], rj[c]))
    sgd, dmins, dmax = op.inisgdm(dev(z["x"]))
    rs, rd, rdm = oracle.inisgdm(z["x"])
    for c in range(7):
        report(f"sgd{c}", normwise_rel(host(sgd)[c], rs[c]))
    report("dmins", normwise_rel(host(dmins), rd))
    assert dmax == pytest.approx(rdm, rel=1e-15)


def test_jacobeano_degenerate_grid_is_an_error():
    """The reference STOPs on |J| <= 1e-20 (solver/jacob_malla.f90:136-139);
    the library returns EMX_ERR_METRIC."""
    from emix3d_amd._lib import EmxError
    shape = (8, 8, 8)
    x = fields.grid_uniform(*shape).numpy().copy()
    x[2] = 0.0  # flat in z: J = 0
    op = Operator(*shape, "c4")
    with pytest.raises(EmxError) as e:
        op.jacobeano(dev(x))
    assert e.value.status == -4


def stage_inputs(shape, seed, scheme, oracle):
    z = inputs(shape, seed)
    s = SID[scheme]
    nx, ny, nz = shape
    z["jbn"] = oracle.jacobeano(s, z["x"])
    z["dvel"], z["dtemp"] = oracle.deriv_vel(s, z["u"], z["temp"],
                                             (nx - 1.0, ny - 1.0, nz - 1.0))
    z["sgd"], z["dmins"], z["deltamax"] = oracle.inisgdm(z["x"])
    z["c4"] = 1.0 / fields.REYNOLDS
    return z


def test_sgdm(oracle):
    shape = (26, 22, 20)
    z = stage_inputs(shape, 21, "c6c", oracle)
    op = Operator(*shape, "c6c")
    amut, vis, vis5 = op.viscosidad_t(dev(z["u"]), dev(z["dvel"]), dev(z["jbn"]), dev(z["sgd"]),
                                      z["c4"])
    ra, rv, r5 = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    assert (ra == 0).any() and (ra > 0).any(), "selective switch not exercised"
    report("amut", normwise_rel(host(amut), ra))
    report("vis", normwise_rel(host(vis), rv))
    report("vis5", normwise_rel(host(vis5), r5))
    assert np.array_equal(host(amut) == 0.0, ra == 0.0), "switch decisions differ"


@pytest.mark.parametrize("scheme", SCHEMES)
def test_flujos_divergencia(oracle, scheme):
    shape = (22, 18, 20)
    z = stage_inputs(shape, 31, scheme, oracle)
    ra, vis, _ = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    op = Operator(*shape, scheme)
    ih = op.default_ih()
    D = {k: dev(z[k]) for k in ("um", "u", "pres", "jbn", "dvel")}
    D["vis"] = dev(vis)
    for neq in range(1, 6):
        rflux = oracle.flujos(neq, z["um"], z["u"], z["pres"], vis, z["jbn"], z["dvel"])
        flux = op.flujos(neq, D["um"], D["u"], D["pres"], D["vis"], D["jbn"], D["dvel"])
        report(f"flux{neq}", normwise_rel(host(flux), rflux))
        rrs, rrsv = oracle.divergencia(SID[scheme], rflux, z["jbn"], ih)
        rs, rsv = op.divergencia(dev(rflux), D["jbn"])
        report(f"rs{neq}", normwise_rel(host(rs), rrs))
        report(f"rsv{neq}", normwise_rel(host(rsv), rrsv))
        rs2, rsv2 = op.flujos_divergencia(neq, D["um"], D["u"], D["pres"], D["vis"], D["jbn"],
                                          D["dvel"])
        report(f"fused rs{neq}", normwise_rel(host(rs2), rrs))
        report(f"fused rsv{neq}", normwise_rel(host(rsv2), rrsv))


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("plan", ["default", "strided", "wave_x_strided16", "strided8",
                                  "materialized_all", "materialized_x"])
def test_flux_div_all_equations(oracle, scheme, plan):
    """The RUNGEK equation loop in three passes equals flujos+divergencia of
    each equation, for every grouping the launch shapes select."""
    shape = (26, 21, 19)
    z = stage_inputs(shape, 33, scheme, oracle)
    _, vis, _ = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    op = Operator(*shape, scheme)
    if plan == "strided":
        op.set_plan(0, 2, 8, op=1)
        op.set_plan(1, 2, 4, op=1)
        op.set_plan(2, 3, 8, op=1)
    elif plan == "wave_x_strided16":
        op.set_plan(0, 0, 2, op=1)
        op.set_plan(1, 2, 8, op=1)
        op.set_plan(2, 3, 4, op=1)
    elif plan == "strided8":
        op.set_plan(0, 4, 4, op=1)
        op.set_plan(1, 4, 2, op=1)
        op.set_plan(2, 4, 4, op=1)
    elif plan == "materialized_all":
        for d in range(3):
            op.set_plan(d, 5, 0, op=1)
    elif plan == "materialized_x":
        # xi first from scratch, then the fused eta and zeta passes
        op.set_plan(0, 5, 0, op=1)
        op.set_plan(1, 2, 4, op=1)
        op.set_plan(2, 0, 2, op=1)
    ih = op.default_ih()
    D = {k: dev(z[k]) for k in ("um", "u", "pres", "jbn", "dvel")}
    rs, rsv = op.flujos_divergencia_all(D["um"], D["u"], D["pres"], dev(vis), D["jbn"],
                                        D["dvel"])
    rs, rsv = host(rs), host(rsv)
    for neq in range(1, 6):
        flux = oracle.flujos(neq, z["um"], z["u"], z["pres"], vis, z["jbn"], z["dvel"])
        rrs, rrsv = oracle.divergencia(SID[scheme], flux, z["jbn"], ih)
        report(f"all rs{neq}", normwise_rel(rs[neq - 1], rrs))
        report(f"all rsv{neq}", normwise_rel(rsv[neq - 1], rrsv))


def sphere_solid(shape, seed):
    x = fields.grid_uniform(*shape)
    phi = fields.level_set_sphere_surface(x, seed)
    return -phi.numpy().copy()  # solid (sobject > 0) inside the sphere/liquid region


@pytest.mark.parametrize("shape", [(12, 10, 9), (30, 20, 25), (60, 60, 60)])
def test_ls_extrapolate(oracle, shape):
    z = inputs(shape, 41)
    nx, ny, nz = shape
    jbn = oracle.jacobeano(2, z["x"])
    sob = sphere_solid(shape, 3)
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    _, dmins, _ = oracle.inisgdm(z["x"])
    dmin = float(dmins.min())
    ref = oracle.ls_extrapolate(z["pres"], jbn, sob, dxyz, dmin)
    op = Operator(*shape, "c6c")
    p = dev(z["pres"])
    op.advect_level_set(p, dev(jbn), dev(sob), dxyz, dmin)
    got = host(p)
    report(f"pres {shape}", normwise_rel(got, ref))
    assert np.array_equal(got[sob <= 0], z["pres"][sob <= 0]) or \
        normwise_rel(got[sob <= 0], z["pres"][sob <= 0]) < 1e-15


@pytest.mark.parametrize("shape", [(12, 10, 9), (32, 24, 20), (48, 40, 36)])
def test_ls_advect_weno5(oracle, shape):
    nx, ny, nz = shape
    z = inputs(shape, 51, curv=True, turb=False)
    jbn = oracle.jacobeano(2, z["x"])
    phi = fields.level_set_sphere_surface(torch.from_numpy(z["x"]), 9).numpy().copy()
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    umax = np.abs(z["u"]).max()
    dt = 0.5 * min(dxyz) / umax * 0.05
    ref = oracle.ls_advect_weno5(phi, z["u"], jbn, dxyz, dt)
    op = Operator(*shape, "c6c")
    p = dev(phi)
    op.advect_weno5(p, dev(z["u"]), dev(jbn), dxyz, dt)
    report(f"phi {shape}", normwise_rel(host(p), ref))
    assert normwise_rel(ref, phi) > 1e-8  # the step moved phi


def test_reference_drop_in_symbols(oracle):
    """cudaDeriv / cuDivergence / advect_CUDA with host arrays, called with
    the arguments the reference Fortran passes (1/deltax spacings)."""
    from emix3d_amd import _lib
    L = _lib.lib()
    shape = (20, 14, 16)
    nx, ny, nz = shape
    z = stage_inputs(shape, 61, "c6c", oracle)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    dvel = np.empty((9, nz, ny, nx))
    dtemp = np.empty((3, nz, ny, nx))
    dummy = np.zeros((nz, ny, nx))
    L.cudaDeriv(P(dvel), P(dtemp), P(z["u"]), P(z["temp"]), P(dummy), P(dummy), P(z["jbn"]),
                1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1), nx, ny, nz)
    ih = (1.0 / (1.0 / (nx - 1)), 1.0 / (1.0 / (ny - 1)), 1.0 / (1.0 / (nz - 1)))
    rdv, rdt = oracle.deriv_vel(2, z["u"], z["temp"], ih)
    report("cudaDeriv dvel", normwise_rel(dvel, rdv))
    report("cudaDeriv dtemp", normwise_rel(dtemp, rdt))
    _, vis, _ = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    flux = oracle.flujos(3, z["um"], z["u"], z["pres"], vis, z["jbn"], z["dvel"])
    keep = flux.copy()
    rs, rsv = np.empty((nz, ny, nx)), np.empty((nz, ny, nx))
    fl = [np.ascontiguousarray(flux[q]) for q in range(6)]
    L.cuDivergence(P(rs), P(rsv), *[P(a) for a in fl], P(z["jbn"]), 1.0 / (nx - 1),
                   1.0 / (ny - 1), 1.0 / (nz - 1), nx, ny, nz)
    rrs, rrsv = oracle.divergencia(2, keep, z["jbn"], ih)
    report("cuDivergence rs", normwise_rel(rs, rrs))
    report("cuDivergence rsv", normwise_rel(rsv, rrsv))
    assert np.array_equal(np.stack(fl), keep)
    sob = sphere_solid(shape, 4)
    pres = z["pres"].copy()
    dmin = float(z["dmins"].min())
    L.advect_CUDA(P(pres), P(z["jbn"]), P(sob), 1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1),
                  nx, ny, nz, dmin)
    rp = oracle.ls_extrapolate(z["pres"], z["jbn"], sob, (1.0 / (nx - 1), 1.0 / (ny - 1),
                                                          1.0 / (nz - 1)), dmin)
    report("advect_CUDA pres", normwise_rel(pres, rp))


def test_host_flujos_and_sgdm(oracle):
    from emix3d_amd import _lib
    L = _lib.lib()
    shape = (12, 11, 10)
    nx, ny, nz = shape
    z = stage_inputs(shape, 71, "c6c", oracle)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    amut, vis, vis5 = (np.empty((nz, ny, nx)) for _ in range(3))
    assert L.emx_host_sgdm(P(z["u"]), P(z["dvel"]), P(z["jbn"]), P(z["sgd"]), z["c4"], P(amut),
                           P(vis), P(vis5), nx, ny, nz) == 0
    ra, rv, r5 = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    report("host amut", normwise_rel(amut, ra))
    report("host vis", normwise_rel(vis, rv))
    for neq in range(1, 6):
        out = [np.empty((nz, ny, nx)) for _ in range(6)]
        assert L.emx_host_flujos(neq, P(z["um"]), P(z["u"]), P(z["pres"]), P(rv), P(z["jbn"]),
                                 P(z["dvel"]), *[P(a) for a in out], nx, ny, nz) == 0
        ref = oracle.flujos(neq, z["um"], z["u"], z["pres"], rv, z["jbn"], z["dvel"])
        report(f"host flux{neq}", normwise_rel(np.stack(out), ref))


def test_full_size_polynomial_exactness():
    """At the 256^3 benchmark size: all three schemes differentiate a
    quartic (C4, C6C) or cubic (C6F) exactly on a uniform grid, in every
    direction, so every line of the full grid is checked without the oracle."""
    n = 256
    op = {s: Operator(n, n, n, s) for s in SCHEMES}
    s = torch.linspace(0.0, 1.0, n, dtype=torch.float64, device="cuda")
    for d in range(3):
        shape = [1, 1, 1]
        shape[2 - d] = n
        xi = s.view(shape).expand(n, n, n)
        for scheme, deg in (("c4", 4), ("c6f", 3), ("c6c", 4)):
            f = (xi ** deg - 0.5 * xi ** 2 + 0.25 * xi).contiguous()
            exact = (deg * xi ** (deg - 1) - xi + 0.25)
            got = op[scheme].derivative(f, d, float(n - 1))
            err = (got - exact).abs().max().item() / exact.abs().max().item()
            report(f"{scheme} x^{deg} dir{d} 256^3", err)
    torch.cuda.synchronize()


def test_full_size_lines_against_oracle(oracle):
    """256^3 random field: 64 random lines per direction against the oracle's
    line solve (the full field is too large for the serial oracle)."""
    n = 256
    rng = np.random.default_rng(99)
    f = torch.from_numpy(rng.standard_normal((n, n, n))).cuda()
    op = Operator(n, n, n, "c6c")
    for d in range(3):
        got = host(op.derivative(f, d, float(n - 1)))
        fh = host(f)
        for _ in range(64):
            a, b = rng.integers(0, n, size=2)
            if d == 0:
                line, gl = fh[a, b, :], got[a, b, :]
            elif d == 1:
                line, gl = fh[a, :, b], got[a, :, b]
            else:
                line, gl = fh[:, a, b], got[:, a, b]
            ref = oracle.deriv_line(2, line, float(n - 1))
            assert normwise_rel(gl, ref) <= TOL
    print("  256^3 sampled lines ok", flush=True)


def test_rk_update_sub_steps(oracle):
    """The three SSP-RK3 sub-step updates against the oracle (chained, each
    sub-step fed the previous GPU result), in place (um1 aliasing um), and
    the argument checks."""
    from emix3d_amd._lib import EmxError
    shape = (37, 19, 23)
    nx, ny, nz = shape
    rng = np.random.default_rng(91)
    um0 = rng.standard_normal((5, nz, ny, nx))
    rs = [rng.standard_normal((5, nz, ny, nx)) for _ in range(3)]
    rsv = [rng.standard_normal((5, nz, ny, nx)) for _ in range(3)]
    op = Operator(*shape, "c6c")
    dt = 3.7e-3
    um = None
    for sub in (1, 2, 3):
        got = host(op.rk_update(sub, dev(um0), None if um is None else dev(um), dev(rs[sub - 1]),
                                dev(rsv[sub - 1]), dt))
        ref = oracle.rk_update(sub, um0, um, rs[sub - 1], rsv[sub - 1], dt)
        report(f"rk sub-step {sub}", normwise_rel(got, ref))
        um = got
    inplace = dev(um)
    op.rk_update(3, dev(um0), inplace, dev(rs[2]), dev(rsv[2]), dt, um1=inplace)
    report("rk in place", normwise_rel(host(inplace), oracle.rk_update(3, um0, um, rs[2], rsv[2],
                                                                       dt)))
    for sub, u in ((0, None), (4, dev(um)), (2, None)):
        with pytest.raises(EmxError):
            op.rk_update(sub, dev(um0), u, dev(rs[0]), dev(rsv[0]), dt)


def test_timing_select_records_one_label():
    """Per-launch timing restricted to one kernel label records only that
    label; without a label every launch is recorded."""
    n = 40
    op = Operator(n, n, n, "c6c")
    f = torch.randn((n, n, n), dtype=torch.float64, device="cuda")
    um = torch.randn((5, n, n, n), dtype=torch.float64, device="cuda")
    op.timing_begin(label="deriv")
    op.derivative(f, 1)
    op.rk_update(1, um, None, um, um, 0.1)
    ents = op.timing_end()
    assert [e["label"] for e in ents] == ["deriv"] and ents[0]["launches"] == 1
    op.timing_begin()
    op.derivative(f, 1)
    op.rk_update(1, um, None, um, um, 0.1)
    assert sorted(e["label"] for e in op.timing_end()) == ["deriv", "rk_update"]


@pytest.mark.parametrize("shape", [(40, 9, 8), (100, 7, 6), (130, 6, 5), (300, 6, 5),
                                   (700, 5, 5)])
def test_flux_div_all_long_xi_lines(oracle, shape):
    """The projection fused with the xi solves for every xi chunk length the
    wavefront kind takes (C = 1, 2, 4, 8, 16; the last needs more than 64 KB
    of LDS and keeps no prefetch), against the oracle."""
    z = stage_inputs(shape, 37, "c6c", oracle)
    _, vis, _ = oracle.sgdm(z["u"], z["dvel"], z["jbn"], z["sgd"], z["c4"])
    op = Operator(*shape, "c6c")
    ih = op.default_ih()
    D = {k: dev(z[k]) for k in ("um", "u", "pres", "jbn", "dvel")}
    rs, rsv = op.flujos_divergencia_all(D["um"], D["u"], D["pres"], dev(vis), D["jbn"],
                                        D["dvel"])
    rs, rsv = host(rs), host(rsv)
    for neq in range(1, 6):
        flux = oracle.flujos(neq, z["um"], z["u"], z["pres"], vis, z["jbn"], z["dvel"])
        rrs, rrsv = oracle.divergencia(2, flux, z["jbn"], ih)
        report(f"{shape} rs{neq}", normwise_rel(rs[neq - 1], rrs))
        report(f"{shape} rsv{neq}", normwise_rel(rsv[neq - 1], rrsv))
