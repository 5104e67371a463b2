"""The device entry points inside a HIP graph (torch.cuda.CUDAGraph on
ROCm): the RK-stage operator (emx_grad, emx_sgdm, emx_flux_div_all), the
dynamic Smagorinsky model (emx_dynsmag) and the learned closure's inference
(emx_mlp_nut) are captured once on a 36x32x28 curvilinear grid and replayed
three times, each replay after eager work that overwrites every output with
NaN and a host synchronisation; every replay equals the eager call bit for
bit (the same kernels on the same inputs)."""
import pytest
import torch

from emix3d_amd import Operator, fields

pytestmark = pytest.mark.gpu


def capture(fn):
    """Warm up fn on a side stream, then capture it into a graph."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def replay_and_check(g, outs, ref, name):
    for r in range(3):
        for t in outs.values():
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in ref:
            assert torch.equal(outs[k], ref[k]), (name, r, k)
    print(f"  {name}: 3 replays equal the eager call", flush=True)


def test_stage_and_closures_in_a_graph():
    shape = (36, 32, 28)
    f = fields.stage_fields(*shape, 77, curvilinear=True, turbulent=True, device="cuda")
    op = Operator(*shape, "c6f")
    jbn = op.jacobeano(f["x"])
    sgd, _, _ = op.inisgdm(f["x"])
    c4 = 1.0 / fields.REYNOLDS
    args = (f["um"], f["u"], f["temp"], f["pres"], jbn, sgd, c4)

    # the stage
    eager = op.rk_stage(*args)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in eager.items()}
    assert all(torch.isfinite(v).all() for v in ref.values())
    work = op.stage_buffers()
    g = capture(lambda: op.rk_stage(*args, work=work))
    replay_and_check(g, work, ref, "rk_stage")

    # dynamic Smagorinsky, then the network, on the stage's gradient
    dvel, delta = ref["dvel"], sgd[0].contiguous()
    ref_d = {"nut": op._empty(), "vis": op._empty()}
    op.dynsmag(f["u"], dvel, jbn, delta, nut=ref_d["nut"], c4=c4, vis=ref_d["vis"])
    torch.cuda.synchronize()
    outs = {"nut": op._empty(), "vis": op._empty()}
    g = capture(lambda: op.dynsmag(f["u"], dvel, jbn, delta, nut=outs["nut"], c4=c4,
                                   vis=outs["vis"]))
    replay_and_check(g, outs, ref_d, "dynsmag")

    from test_oracle_mlp import rand_norm, rand_params
    params = torch.from_numpy(rand_params(31)).cuda()
    norm = torch.from_numpy(rand_norm(32)).cuda()
    ref_m = {"nut": op._empty(), "vis": op._empty()}
    op.mlp_nut(params, norm, dvel, jbn, delta, nut=ref_m["nut"], c4=c4, vis=ref_m["vis"])
    torch.cuda.synchronize()
    assert torch.isfinite(ref_m["nut"]).all()
    outs_m = {"nut": op._empty(), "vis": op._empty()}
    g = capture(lambda: op.mlp_nut(params, norm, dvel, jbn, delta, nut=outs_m["nut"], c4=c4,
                                   vis=outs_m["vis"]))
    replay_and_check(g, outs_m, ref_m, "mlp_nut")
