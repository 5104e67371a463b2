"""The operator's five workloads at their stated sizes, against the oracle.

  * 64^3 uniform grid, 8 sine modes per axis and component: the C6F (and
    C4, C6C) gradient fused with the velocity divergence (emx_grad_div)
    against orc_deriv_vel + orc_divergencia (velocity as the flux);
  * 256^3 tanh-stretched grid (beta = 2), solenoidal k^-5/3 field plus the
    six-blade jet: one whole RK stage (DERIV_VEL, viscosidad_t, viscosidad,
    five flujos + divergencia; solver/rungek_ico.f90:79-95) against the
    oracle's composition, every output at 1e-12 normwise relative; the
    metrics and SGS geometry of the grid against JACOBEANO and INISGDM;
  * the batched stage's 48 boxes (32^3 .. 192x128x96, 25.7 M points):
    every box's gradient, divergence and WENO5-RK3 level-set step against
    orc_deriv_vel, orc_divergencia and orc_ls_advect_weno5, after the
    autotuner and from the HIP graph;
  * adjoint at 128^3 stretched: see test_gpu_adjoint.py (dot test, finite
    differences and the oracle's transpose at that size);
  * BF16 matrix-core inference of the trained closure on the 256^3 stage
    field against the oracle's FP64 network on 2^18 sampled points (3e-2:
    BF16 operands keep 8 significant bits).
Oracle calls run in a thread pool (the C oracle holds no global state and
ctypes releases the interpreter lock).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Batch, Operator, fields

pytestmark = pytest.mark.gpu
TOL = 1e-12
C4, C6F, C6C = 0, 1, 2
SCHEMES = {"c4": C4, "c6f": C6F, "c6c": C6C}


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def report(name, err, tol=TOL):
    print(f"  {name}: {err:.2e}", flush=True)
    assert err <= tol, (name, err)


def pool():
    return ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1))


@pytest.mark.parametrize("scheme", ["c6f", "c4", "c6c"])
def test_sine_modes_64_grad_div(oracle, scheme):
    n = 64
    x = fields.grid_uniform(n, n, n, device="cuda")
    u = fields.velocity_sines(x, 20261203).contiguous()
    op = Operator(n, n, n, scheme)
    jbn = op.jacobeano(x)
    dvel, div = op.grad_div(u, jbn=jbn)
    hu, hj = host(u), host(jbn)
    ih = (n - 1.0,) * 3
    s = SCHEMES[scheme]
    rdv, _ = oracle.deriv_vel(s, hu, hu[0].copy(), ih)
    rdiv, _ = oracle.divergencia(s, np.concatenate([hu, np.zeros_like(hu)]), hj, ih)
    report(f"{scheme} 64^3 dvel", normwise_rel(host(dvel), rdv))
    report(f"{scheme} 64^3 div", normwise_rel(host(div), rdiv))
    report(f"{scheme} 64^3 jbn", normwise_rel(hj, oracle.jacobeano(s, host(x))))


@pytest.mark.timeout(600)
def test_stage_256_curvilinear(oracle):
    n = 256
    f = fields.stage_fields(n, n, n, 20261204, curvilinear=True, turbulent=True, device="cuda")
    op = Operator(n, n, n, "c6f")
    jbn = op.jacobeano(f["x"])
    sgd, _, _ = op.inisgdm(f["x"])
    c4 = 1.0 / fields.REYNOLDS
    w = op.rk_stage(f["um"], f["u"], f["temp"], f["pres"], jbn, sgd, c4)
    g = {k: host(v) for k, v in w.items()}
    z = {k: host(v) for k, v in f.items()}
    hj, hs = host(jbn), host(sgd)
    del f, w
    torch.cuda.empty_cache()
    ih = (n - 1.0,) * 3
    with pool() as ex:
        fj = ex.submit(oracle.jacobeano, C6F, z["x"])
        fs = ex.submit(oracle.inisgdm, z["x"])
        # the stage is fed the GPU's metrics and SGS geometry (checked
        # against the oracle's above), as the Fortran path receives them
        rdv, rdt = oracle.deriv_vel(C6F, z["u"], z["temp"], ih)
        amut, vis, _ = oracle.sgdm(z["u"], rdv, hj, hs, c4)

        def eq(neq):
            flux = oracle.flujos(neq, z["um"], z["u"], z["pres"], vis, hj, rdv)
            return oracle.divergencia(C6F, flux, hj, ih)

        res = list(ex.map(eq, range(1, 6)))
        report("256^3 jbn (JACOBEANO)", normwise_rel(hj, fj.result()))
        report("256^3 sgd (INISGDM)", normwise_rel(hs, fs.result()[0]))
    report("256^3 dvel", normwise_rel(g["dvel"], rdv))
    report("256^3 dtemp", normwise_rel(g["dtemp"], rdt))
    report("256^3 amut", normwise_rel(g["amut"], amut))
    report("256^3 vis", normwise_rel(g["vis"], vis))
    for m, (rs, rsv) in enumerate(res):
        report(f"256^3 rs{m + 1}", normwise_rel(g["rs"][m], rs))
        if m in (1, 2, 3):
            report(f"256^3 rsv{m + 1}", normwise_rel(g["rsv"][m], rsv))
        else:
            assert not g["rsv"][m].any(), "rsv of continuity / scalar must be zero"
    # the structure-function switch takes the same decisions
    assert np.array_equal(g["amut"] == 0.0, amut == 0.0)


@pytest.mark.timeout(600)
def test_batched_48_boxes(oracle):
    dims = fields.batch_dims(20261205, 48)
    assert len(dims) == 48 and (32, 32, 32) in dims and (192, 128, 96) in dims
    b = Batch(dims, "c6c")
    boxes = []
    for i, d in enumerate(dims):
        z = fields.batch_box(d, 1000 + i, device="cuda")
        z["jbn"] = Operator(*d, "c6c").jacobeano(z["x"])
        del z["x"]
        nx, ny, nz = d
        z["phi0"] = z["phi"].clone()
        z["dvel"] = torch.empty((9, nz, ny, nx), dtype=torch.float64, device="cuda")
        z["div"] = torch.empty((nz, ny, nx), dtype=torch.float64, device="cuda")
        b.bind(i, z["u"], z["jbn"], z["phi"], z["dvel"], z["div"], z["ih"], z["dxyz"], z["dt"])
        boxes.append(z)
    b.tune(reps=2)
    for z in boxes:
        z["phi"].copy_(z["phi0"])
    b.stage(graph=True)
    torch.cuda.synchronize()
    print(f"  {sum(np.prod(d) for d in dims) / 1e6:.1f} M points; plans "
          f"{[b.get_plan(d) for d in range(3)]}", flush=True)

    def check(i):
        z = boxes[i]
        u, jbn = z["u"].cpu().numpy(), z["jbn"].cpu().numpy()
        dvel, _ = oracle.deriv_vel(C6C, u, u[0].copy(), z["ih"])
        div, _ = oracle.divergencia(C6C, np.concatenate([u, np.zeros_like(u)]), jbn, z["ih"])
        phi = oracle.ls_advect_weno5(z["phi0"].cpu().numpy(), u, jbn, z["dxyz"], z["dt"])
        return (normwise_rel(z["dvel"].cpu().numpy(), dvel),
                normwise_rel(z["div"].cpu().numpy(), div),
                normwise_rel(z["phi"].cpu().numpy(), phi))

    with pool() as ex:
        errs = list(ex.map(check, range(len(dims))))
    for name, k in (("dvel", 0), ("div", 1), ("phi", 2)):
        worst = int(np.argmax([e[k] for e in errs]))
        report(f"48 boxes {name} (worst box {worst} {dims[worst]})", errs[worst][k])


@pytest.mark.timeout(600)
@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "emix3d_amd",
                                                    "closure_weights.npz")),
                    reason="no trained closure weights in the tree")
def test_learned_closure_inference_256(oracle):
    from emix3d_amd.closure import LearnedClosure
    n = 256
    f = fields.stage_fields(n, n, n, 20261204, curvilinear=True, turbulent=True, device="cuda")
    op = Operator(n, n, n, "c6f")
    jbn = op.jacobeano(f["x"])
    sgd, _, _ = op.inisgdm(f["x"])
    dvel, _ = op.deriv_vel(f["u"])
    delta = sgd[0].contiguous()
    cl = LearnedClosure.load()
    nut = cl.nut(op, dvel, jbn, delta)
    got = host(nut).ravel()
    assert np.all(got >= 0.0)
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(n ** 3, 1 << 18, replace=False))
    feat = oracle.mlp_features(host(dvel), host(jbn), host(delta))[idx]
    P, norm = cl.params.cpu().numpy(), cl.norm.cpu().numpy()
    ref = oracle.mlp_forward(P, norm, feat)
    report("256^3 bf16 inference (2^18 sampled points)", normwise_rel(got[idx], ref), 3e-2)
