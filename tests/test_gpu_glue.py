"""RUNGEK glue on the GPU (SSP-RK3 update aside, tested with the operator):
boundary values, conversions, velocity clip and rebuild against the oracle
(pinned in test_oracle_glue.py), and one device sub-step against the oracle
composition.  Pointwise glue is compared exactly (the same IEEE operations
in the same order); the sub-step at the operator's 1e-12 bound."""
import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Operator, fields

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def state(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    um1 = rng.standard_normal((5, nz, ny, nx))
    um1[0] = 0.5 + rng.random((nz, ny, nx))
    um1[1:4] *= 8.0
    mask = (rng.random((nz, ny, nx)) > 0.3).astype(np.float64)
    return um1, mask


@pytest.mark.parametrize("shape", [(9, 8, 7), (37, 19, 23), (64, 64, 64), (5, 5, 5)])
def test_glue_kernels_vs_oracle(oracle, shape):
    nx, ny, nz = shape
    um1, mask = state(shape, sum(shape))
    mach = 0.3
    op = Operator(*shape, "c6c")
    print(f"  glue {shape}", flush=True)
    d = dev(um1)
    op.frontera_val(d)
    assert np.array_equal(host(d), oracle.frontera_val(um1)), "frontera_val"
    u, p, t = op.varc_to_varnc(dev(um1), mach)
    ru, rp, rt, _ = oracle.varc_to_varnc(um1, mach)
    for name, a, b in (("u", u, ru), ("pres", p, rp), ("temp", t, rt)):
        assert np.array_equal(host(a), b), name
    um0 = op.varnc_to_varc(dev(ru), dev(rp), dev(rt), mach)
    assert np.array_equal(host(um0), oracle.varnc_to_varc(ru, rp, rt, mach)[0]), "varnc_to_varc"
    d = dev(um1)
    u, p, t, c = (op._empty(3), op._empty(), op._empty(), op._empty())
    op.rk_clip(d, dev(mask), mach, u, p, t, c)
    ro, ru, rp, rt, rc = oracle.rk_clip(um1, mask, mach)
    for name, a, b in (("um1", d, ro), ("u", u, ru), ("pres", p, rp), ("temp", t, rt),
                       ("conc", c, rc)):
        assert np.array_equal(host(a), b), f"clip {name}"
    d = dev(um1)
    op.rk_rebuild(d, dev(ru), dev(rp), mach)
    assert np.array_equal(host(d), oracle.rk_rebuild(um1, ru, rp, mach)), "rebuild"


def test_glue_argument_checks():
    from emix3d_amd._lib import EmxError
    op = Operator(8, 8, 8, "c6c")
    um1 = torch.ones((5, 8, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(EmxError):
        op.varc_to_varnc(um1, 0.0)
    with pytest.raises(EmxError):
        op.rk_rebuild(um1, um1[:3], um1[0], -1.0)


def test_device_substep_vs_oracle(oracle):
    """One RUNGEK sub-step (sub-step 2) on the device against the oracle's
    composition of the same steps, with the pressure extrapolated into a
    spherical solid."""
    shape = (24, 20, 18)
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 77, curvilinear=True, turbulent=False)
    z = {k: np.ascontiguousarray(v.numpy()) for k, v in f.items()}
    sid = 2
    z["jbn"] = oracle.jacobeano(sid, z["x"])
    sgd, dmins, _ = oracle.inisgdm(z["x"])
    c4 = 1.0 / fields.REYNOLDS
    rng = np.random.default_rng(78)
    um0 = z["um"] * (1.0 + 0.01 * rng.standard_normal(z["um"].shape))
    mask = (rng.random((nz, ny, nx)) > 0.1).astype(np.float64)
    xs = fields.grid_uniform(*shape)
    sobject = -np.ascontiguousarray(fields.level_set_sphere_surface(xs, 5).numpy())
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    deltamin = min(dxyz)
    dt, mach, sub = 2e-3, 0.3, 2
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    # oracle composition
    dvel, _ = oracle.deriv_vel(sid, z["u"], z["temp"], ih)
    _, vis, _ = oracle.sgdm(z["u"], dvel, z["jbn"], sgd, c4)
    rs = np.empty_like(z["um"])
    rsv = np.empty_like(z["um"])
    for neq in range(1, 6):
        fl = oracle.flujos(neq, z["um"], z["u"], z["pres"], vis, z["jbn"], dvel)
        rs[neq - 1], rsv[neq - 1] = oracle.divergencia(sid, fl, z["jbn"], ih)
    um1 = oracle.rk_update(sub, um0, z["um"], rs, rsv, dt)
    um1 = oracle.frontera_val(um1)
    um1, u, pres, temp, conc = oracle.rk_clip(um1, mask, mach)
    pres = oracle.ls_extrapolate(pres, z["jbn"], sobject, dxyz, deltamin)
    um1 = oracle.rk_rebuild(um1, u, pres, mach)
    # device
    op = Operator(*shape, "c6c")
    D = {k: dev(z[k]) for k in ("um", "u", "temp", "pres", "jbn")}
    got = op.rk_substep(sub, dev(um0), D["um"], D["u"], D["temp"], D["pres"], D["jbn"],
                        dev(sgd), c4, dt, mach, dev(mask), sobject=dev(sobject), dxyz=dxyz,
                        deltamin=deltamin)
    errs = {"um1": normwise_rel(host(got), um1), "u": normwise_rel(host(D["u"]), u),
            "pres": normwise_rel(host(D["pres"]), pres), "temp": normwise_rel(host(D["temp"]), temp)}
    print("  sub-step", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-12, errs


@pytest.mark.parametrize("radius", [0.0, 0.12, 2.0])
def test_extrapolation_solid_sizes(oracle, radius):
    """Pressure extrapolation over the solid-point list: no solid (radius 0:
    nothing changes), a small central ball, and a solid that covers the
    whole box, against the oracle."""
    shape = (30, 26, 22)
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 81, curvilinear=True, turbulent=False)
    x = np.ascontiguousarray(f["x"].numpy())
    jbn = oracle.jacobeano(2, x)
    pres = np.ascontiguousarray(f["pres"].numpy())
    xs = fields.grid_uniform(*shape).numpy()
    sob = np.ascontiguousarray(radius - np.sqrt(((xs - 0.5) ** 2).sum(0)))
    if radius == 0.0:
        sob = -np.abs(sob) - 1e-3
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    ref = oracle.ls_extrapolate(pres, jbn, sob, dxyz, min(dxyz))
    op = Operator(*shape, "c6c")
    p = dev(pres)
    op.advect_level_set(p, dev(jbn), dev(sob), dxyz, min(dxyz))
    got = host(p)
    print(f"  radius {radius}: solid points {(sob > 0).sum()}, "
          f"err {normwise_rel(got, ref):.1e}", flush=True)
    assert normwise_rel(got, ref) <= 1e-12
    assert np.array_equal(got[sob <= 0], pres[sob <= 0])
    if radius > 0:
        assert not np.array_equal(got[sob > 0], pres[sob > 0])


def test_substep_graph_replays_match_eager(oracle):
    """A whole device sub-step captured in one HIP graph (torch.cuda.graph,
    the library launching on the capture stream) and replayed three times,
    with eager copies and synchronisations between the replays: every replay
    equals the eager sub-step bit for bit."""
    shape = (40, 36, 32)
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 90, curvilinear=True, turbulent=True, device="cuda")
    op = Operator(*shape, "c6c")
    jbn = op.jacobeano(f["x"])
    sgd, _, _ = op.inisgdm(f["x"])
    c4 = 1.0 / fields.REYNOLDS
    gen = torch.Generator(device="cuda").manual_seed(91)
    um0 = f["um"] * (1.0 + 0.01 * torch.randn(f["um"].shape, dtype=torch.float64,
                                               device="cuda", generator=gen))
    mask = (torch.rand(f["pres"].shape, dtype=torch.float64, device="cuda",
                       generator=gen) > 0.05).double()
    xs = fields.grid_uniform(*shape, device=torch.device("cuda"))
    sob = (0.2 - torch.sqrt(((xs - 0.5) ** 2).sum(0))).contiguous()
    dxyz = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    init = {k: f[k].clone() for k in ("um", "u", "temp", "pres")}
    st = {k: v.clone() for k, v in init.items()}
    work = op.stage_buffers()
    um1 = op._empty(5)

    def run():
        op.rk_substep(2, um0, st["um"], st["u"], st["temp"], st["pres"], jbn, sgd, c4, 2e-3, 0.3,
                      mask, sobject=sob, dxyz=dxyz, deltamin=min(dxyz), work=work, um1=um1)

    def reset():
        for k, v in init.items():
            st[k].copy_(v)

    run()  # eager: builds every table and launch shape before the capture
    torch.cuda.synchronize()
    ref = {"um1": um1.clone(), **{k: st[k].clone() for k in ("u", "temp", "pres")}}
    reset()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            run()
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        reset()
        um1.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(um1, ref["um1"]), rep
        for k in ("u", "temp", "pres"):
            assert torch.equal(st[k], ref[k]), (rep, k)
        # eager work between replays (the sub-step updates u, temp, pres in
        # place, so from the initial state again)
        reset()
        run()
        torch.cuda.synchronize()
        assert torch.equal(um1, ref["um1"])


def test_load_run_feeds_the_operator(tmp_path):
    """A state written in the reference's formats (field, grid, sobject) and
    loaded with load_run reaches the device unchanged: metrics and the
    conservative variables equal the in-memory path's."""
    from emix3d_amd import fortran_io as fio
    shape = (20, 18, 16)
    nx, ny, nz = shape
    f = fields.stage_fields(nx, ny, nz, 95, curvilinear=True, turbulent=False)
    x = f["x"].numpy()
    sob = np.random.default_rng(2).standard_normal((nz, ny, nx))
    fio.write_field(tmp_path / "field_01.001", f["u"].numpy(), f["temp"].numpy(),
                    f["pres"].numpy())
    fio.write_grid(tmp_path / "grid", x)
    fio.write_scalar(tmp_path / "sobject.field", sob)
    op, t = fio.load_run(tmp_path / "field_01.001", tmp_path / "grid", 0.3,
                         sobject=tmp_path / "sobject.field")
    op2 = Operator(*shape, "c6c")
    jbn = op2.jacobeano(dev(x))
    assert torch.equal(t["jbn"], jbn)
    um = op2.varnc_to_varc(dev(f["u"].numpy()), dev(f["pres"].numpy()), dev(f["temp"].numpy()),
                           0.3)
    assert torch.equal(t["um"], um)
    assert np.array_equal(host(t["sobject"]), sob)
