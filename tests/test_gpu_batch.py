"""Batched multi-grid stage (emx_batch): boxes of uneven sizes, one launch
per phase over all boxes, replayed from a HIP graph.

  * every box's gradient, velocity divergence and advected level set against
    the oracle (orc_deriv_vel, orc_divergencia with the velocity as the flux,
    orc_ls_advect_weno5) at 1e-12 normwise relative, for the graph and the
    eager launch, under every built line-solver shape;
  * graph replays, with eager work and synchronisations between them, equal
    the eager result each time;
  * the wavefront kind along x with each box on its own chunk length (1, 2,
    3, 4 and 8 points per lane in one launch), and its shape checks;
  * metrics changed in place after the bind are used by the next stage;
  * the autotuner keeps a built shape and the results do not change;
  * the single-grid entry emx_grad_div gives the same fields.
"""
import numpy as np
import pytest
import torch

from conftest import normwise_rel
from emix3d_amd import Batch, Operator, fields
from emix3d_amd._lib import EmxError

pytestmark = pytest.mark.gpu
TOL = 1e-12
DIMS = [(6, 7, 5), (24, 9, 13), (17, 20, 8), (9, 9, 30), (33, 12, 11)]
SHAPES = [(8, 4), (8, 8), (16, 4), (16, 8), (32, 4), (32, 8), (64, 8), (64, 16)]
WAVE_SHAPES = [(0, 1), (0, 2), (0, 4), (0, 8)]  # one wavefront per line, direction 0


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def report(name, err):
    print(f"  {name}: {err:.2e}", flush=True)
    assert err <= TOL, (name, err)


def make_batch(dims, scheme="c6c"):
    boxes = []
    b = Batch(dims, scheme)
    for i, d in enumerate(dims):
        z = fields.batch_box(d, 100 + i, device="cuda")
        op = Operator(*d, scheme)
        z["jbn"] = op.jacobeano(z["x"])
        z["phi0"] = z["phi"].clone()
        nx, ny, nz = d
        z["dvel"] = torch.empty((9, nz, ny, nx), dtype=torch.float64, device="cuda")
        z["div"] = torch.empty((nz, ny, nx), dtype=torch.float64, device="cuda")
        b.bind(i, z["u"], z["jbn"], z["phi"], z["dvel"], z["div"], z["ih"], z["dxyz"], z["dt"])
        boxes.append(z)
    return b, boxes


def reference(oracle, z, scheme=2):
    u = z["u"].cpu().numpy()
    jbn = z["jbn"].cpu().numpy()
    dvel, _ = oracle.deriv_vel(scheme, u, u[0].copy(), z["ih"])
    flux = np.concatenate([u, np.zeros_like(u)])
    div, _ = oracle.divergencia(scheme, flux, jbn, z["ih"])
    phi = oracle.ls_advect_weno5(z["phi0"].cpu().numpy(), u, jbn, z["dxyz"], z["dt"])
    return dvel, div, phi


def check(oracle, boxes, refs, tag):
    for i, (z, (dvel, div, phi)) in enumerate(zip(boxes, refs)):
        report(f"{tag} box{i} dvel", normwise_rel(host(z["dvel"]), dvel))
        report(f"{tag} box{i} div", normwise_rel(host(z["div"]), div))
        report(f"{tag} box{i} phi", normwise_rel(host(z["phi"]), phi))


def reset(boxes):
    for z in boxes:
        z["phi"].copy_(z["phi0"])


def test_batch_stage_vs_oracle_every_shape(oracle):
    b, boxes = make_batch(DIMS)
    refs = [reference(oracle, z) for z in boxes]
    for lg, c in SHAPES:
        for d in range(3):
            b.set_plan(d, lg, c)
        for graph in (False, True):
            reset(boxes)
            b.stage(graph=graph)
            check(oracle, boxes, refs, f"lg{lg} C{c} graph={graph}")
    for lg, c in WAVE_SHAPES:
        b.set_plan(0, lg, c)
        b.set_plan(1, 16, 8)
        b.set_plan(2, 32, 4)
        for graph in (False, True):
            reset(boxes)
            b.stage(graph=graph)
            check(oracle, boxes, refs, f"wave C{c} graph={graph}")


@pytest.mark.parametrize("scheme,sid", [("c4", 0), ("c6f", 1)])
def test_batch_stage_other_schemes(oracle, scheme, sid):
    """The C4 and C6F instantiations of the batched kernels (the cases above
    run C6C) against the oracle of the same scheme."""
    b, boxes = make_batch(DIMS, scheme)
    refs = [reference(oracle, z, sid) for z in boxes]
    for lg, c in ((16, 4), (32, 8)):
        for d in range(3):
            b.set_plan(d, lg, c)
        for graph in (False, True):
            reset(boxes)
            b.stage(graph=graph)
            check(oracle, boxes, refs, f"{scheme} lg{lg} C{c} graph={graph}")
    b.set_plan(0, 0, 4)
    reset(boxes)
    b.stage(graph=True)
    check(oracle, boxes, refs, f"{scheme} wave C4")


def test_batch_wave_chunk_per_box(oracle):
    # x lines of 20, 70, 130, 200 and 260 points: 1, 2, 3, 4 and 8 points per
    # lane in one launch of the wavefront kind (each box its shortest chunk)
    dims = [(70, 9, 8), (260, 6, 5), (20, 8, 7), (130, 7, 6), (200, 6, 5)]
    b, boxes = make_batch(dims)
    refs = [reference(oracle, z) for z in boxes]
    assert b.get_plan(0) == (0, 8)  # the only wavefront shape that fits 260
    with pytest.raises(EmxError):
        b.set_plan(0, 0, 4)
    b.set_plan(1, 16, 4)
    b.set_plan(2, 16, 4)
    for graph in (False, True):
        reset(boxes)
        b.stage(graph=graph)
        check(oracle, boxes, refs, f"wave per-box chunks graph={graph}")


def test_batch_wave_shape_only_along_x():
    b, _ = make_batch(DIMS[:2])
    assert b.get_plan(0) == (0, 1)  # default: the shortest chunk that fits
    for d in (1, 2):
        with pytest.raises(EmxError):
            b.set_plan(d, 0, 4)
    with pytest.raises(EmxError):
        b.set_plan(0, 0, 3)  # not built


def test_batch_reads_metrics_in_place(oracle):
    # the metrics are read at every stage, like the velocity: changed in
    # place after the bind, the next stage uses the new values
    b, boxes = make_batch(DIMS[:3])
    reset(boxes)
    b.stage(graph=False)
    for z in boxes:
        z["jbn"][[0, 4, 8]] *= 1.5  # metrics 1, 5, 9
        z["jbn"][9] *= 1.5          # and J
    refs = [reference(oracle, z) for z in boxes]
    for graph in (False, True):
        reset(boxes)
        b.stage(graph=graph)
        check(oracle, boxes, refs, f"metrics changed in place graph={graph}")


def test_graph_replays_match_eager(oracle):
    b, boxes = make_batch(DIMS)
    reset(boxes)
    b.stage(graph=False)
    eager = [(host(z["dvel"]), host(z["div"]), host(z["phi"])) for z in boxes]
    for rep in range(4):
        reset(boxes)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        b.stage(graph=True)
        for z, (dv, dd, ph) in zip(boxes, eager):
            assert np.array_equal(host(z["dvel"]), dv), rep
            assert np.array_equal(host(z["div"]), dd), rep
            assert np.array_equal(host(z["phi"]), ph), rep
        del junk


def test_autotuner_keeps_results(oracle):
    b, boxes = make_batch(DIMS)
    refs = [reference(oracle, z) for z in boxes]
    b.tune(reps=3)
    for d in range(3):
        assert b.get_plan(d) in SHAPES + (WAVE_SHAPES if d == 0 else [])
    reset(boxes)
    b.stage(graph=True)
    check(oracle, boxes, refs, "tuned")


def test_single_grid_grad_div(oracle):
    d = (29, 22, 17)
    z = fields.batch_box(d, 7, device="cuda")
    op = Operator(*d, "c6c")
    jbn = op.jacobeano(z["x"])
    dvel, div = op.grad_div(z["u"], jbn=jbn)
    z["jbn"] = jbn
    z["phi0"] = z["phi"]
    rdv, rdiv, _ = reference(oracle, z)
    report("grad_div dvel", normwise_rel(host(dvel), rdv))
    report("grad_div div", normwise_rel(host(div), rdiv))


def test_rebind_and_replan_without_host_sync(oracle):
    """stage, then bind one box with a new time step and change a line
    shape, then stage again, with no host synchronisation in between: the
    second stage runs on the new descriptors, which are uploaded only once
    the first stage has finished, and every box matches two oracle steps."""
    b, boxes = make_batch(DIMS[:3])
    reset(boxes)
    b.set_plan(1, 16, 4)
    b.stage(graph=True)          # queued, not waited for
    z = boxes[1]
    dt2 = 0.5 * z["dt"]
    b.bind(1, z["u"], z["jbn"], z["phi"], z["dvel"], z["div"], z["ih"], z["dxyz"], dt2)
    b.set_plan(1, 32, 8)
    b.stage(graph=True)
    for i, z in enumerate(boxes):
        dvel, div, phi1 = reference(oracle, z)
        dt = dt2 if i == 1 else z["dt"]
        phi2 = oracle.ls_advect_weno5(phi1, z["u"].cpu().numpy(), z["jbn"].cpu().numpy(),
                                      z["dxyz"], dt)
        report(f"box{i} dvel", normwise_rel(host(z["dvel"]), dvel))
        report(f"box{i} div", normwise_rel(host(z["div"]), div))
        report(f"box{i} phi after two stages", normwise_rel(host(z["phi"]), phi2))
