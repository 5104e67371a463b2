"""Dynamic Smagorinsky and the learned closure on the GPU.

  * emx_dynsmag against the oracle (1e-12 normwise relative);
  * the FP32 training kernels against the oracle's FP64 network: forward
    1e-5, parameter gradient 1e-4 normwise (FP32 accumulation over the
    batch);
  * BF16 matrix-core inference against the oracle's FP64 forward on the same
    features: 3e-2 normwise (BF16 operands carry 8 significant bits);
  * training by the library's own kernels lowers the loss, and the committed
    trained network is closer to dynamic Smagorinsky than the static
    Smagorinsky model (one least-squares constant) and the
    structure-function sgdm on held-out snapshots (what the network gains
    over the models without it).
"""
import os

import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Operator, fields
from emix3d_amd.api import mlp_backward_f32, mlp_forward_f32

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def report(name, err, tol):
    print(f"  {name}: {err:.2e} (tol {tol:.0e})", flush=True)
    assert err <= tol, (name, err)


def small_state(shape, seed, oracle, scheme=2):
    nx, ny, nz = shape
    x = fields.grid_tanh(nx, ny, nz)
    u = fields.velocity_sines(x, seed).numpy().copy()
    x = x.numpy().copy()
    jbn = oracle.jacobeano(scheme, x)
    ih = (nx - 1.0, ny - 1.0, nz - 1.0)
    dvel, _ = oracle.deriv_vel(scheme, u, u[0].copy(), ih)
    sgd, _, _ = oracle.inisgdm(x)
    return {"x": x, "u": u, "jbn": jbn, "dvel": dvel, "delta": np.ascontiguousarray(sgd[0])}


def test_dynsmag_vs_oracle(oracle):
    shape = (20, 18, 16)
    z = small_state(shape, 7, oracle)
    N = np.prod(shape)
    gp = np.empty((9,) + z["u"].shape[1:])
    m = (z["jbn"][0], z["jbn"][4], z["jbn"][8])
    for a in range(3):
        for b in range(3):
            gp[3 * a + b] = m[b] * z["dvel"][3 * b + a]
    ref = np.empty(z["u"].shape[1:])
    from oracle.pyoracle import _check, _ptr, lib
    _check(lib().orc_dynsmag(*shape, _ptr(z["u"]), _ptr(gp), _ptr(z["delta"]), _ptr(ref)),
           "dynsmag")
    assert (ref > 0).sum() > N // 10 and (ref == 0).sum() > 0, "clipping not exercised"
    op = Operator(*shape, "c6c")
    got = host(op.dynsmag(dev(z["u"]), dev(z["dvel"]), dev(z["jbn"]), dev(z["delta"])))
    report("dynsmag nu_t", normwise_rel(got, ref), 1e-12)


def rand_net(seed):
    from test_oracle_mlp import rand_norm, rand_params
    return rand_params(seed), rand_norm(seed + 1)


def test_training_kernels_vs_oracle(oracle):
    P, norm = rand_net(11)
    rng = np.random.default_rng(12)
    n = 2048
    feat = rng.standard_normal((n, 10)).astype(np.float32)
    dnu = rng.standard_normal(n)
    Pd, nd = torch.from_numpy(P).cuda(), torch.from_numpy(norm).cuda()
    fd = torch.from_numpy(feat).cuda()
    act, nu = mlp_forward_f32(Pd, nd, fd)
    ref = oracle.mlp_forward(P, norm, feat.astype(np.float64))
    report("forward f32", normwise_rel(host(nu), ref), 1e-5)
    grad, _ = mlp_backward_f32(Pd, nd, fd, act, torch.from_numpy(dnu.astype(np.float32)).cuda())
    gref = oracle.mlp_grad(P, norm, feat.astype(np.float64), dnu.astype(np.float32).astype(np.float64))
    report("parameter gradient f32", normwise_rel(host(grad), gref), 1e-4)


def test_bf16_inference_vs_oracle(oracle):
    shape = (33, 17, 19)  # N not a multiple of the 16-point tile
    z = small_state(shape, 9, oracle)
    feat = oracle.mlp_features(z["dvel"], z["jbn"], z["delta"])
    mean, std = feat.mean(0), feat.std(0)
    P, _ = rand_net(21)
    norm = np.concatenate([mean, 1.0 / std, [1.3]]).astype(np.float32)
    ref = oracle.mlp_forward(P, norm, feat).reshape(shape[::-1])
    op = Operator(*shape, "c6c")
    got = host(op.mlp_nut(torch.from_numpy(P).cuda(), torch.from_numpy(norm).cuda(),
                          dev(z["dvel"]), dev(z["jbn"]), dev(z["delta"])))
    report("bf16 mfma inference", normwise_rel(got, ref), 3e-2)
    assert np.all(got >= 0.0)


def test_training_lowers_loss():
    from emix3d_amd.closure import LearnedClosure, SnapshotFamily
    fam = SnapshotFamily(48, seeds=range(1, 21))
    cl = LearnedClosure()
    hist = cl.train(fam, steps=120, batch=16384, lr=3e-3)
    first = np.mean([h[0] for h in hist[:10]])
    last = np.mean([h[0] for h in hist[-10:]])
    print(f"  nu_t loss {first:.3e} -> {last:.3e}; flux loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}")
    assert last < 0.75 * first


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "emix3d_amd",
                                                    "closure_weights.npz")),
                    reason="no trained closure weights in the tree")
def test_trained_closure_beats_static_smagorinsky_and_sgdm():
    """Held-out snapshots (sine family seeds 201+, turbulent family seeds
    1101+): the network's nu_t and viscous flux divergence are closer to
    dynamic Smagorinsky than the static Smagorinsky model with the constant
    fitted by least squares to the same training targets, and than the
    structure-function sgdm."""
    from emix3d_amd.closure import LearnedClosure, SnapshotFamily, StaticSmagorinsky, evaluate
    cl = LearnedClosure.load()
    assert cl.static_c is not None and cl.static_c > 0.0
    static = StaticSmagorinsky(cl.static_c)
    for kind, seeds in (("sines", range(201, 203)), ("turbulent", range(1101, 1103))):
        ev = evaluate(SnapshotFamily(128, kind=kind), cl, seeds, static)
        print(f"  held-out {kind} relative L2 vs dynamic Smagorinsky:", ev)
        for q in ("nut", "rsv"):
            assert ev[f"mlp_{q}"] < ev[f"static_{q}"], (kind, q)
            assert ev[f"mlp_{q}"] < ev[f"sgdm_{q}"], (kind, q)
