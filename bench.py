"""Benchmark: grid points per second of one FP64 RK-stage spatial operator.

Workload (the curvilinear full-operator case): a 256^3 FP64 grid with tanh
stretching (beta=2) toward the walls, metrics from JACOBEANO, a seeded
solenoidal k^-5/3 velocity field (k <= 32) plus a six-blade rotating radial
jet at mid-height.  One step = the spatial operator of one RUNGEK sub-step
(solver/rungek_ico.f90:79-95): DERIV_VEL (9 velocity + 3 temperature
gradient components), viscosidad_t + viscosidad (structure-function eddy
viscosity), and flujos + divergencia for the five equations.  Inputs are
resident in HBM when the timed region starts.

Also reported, on the same JSON line:
  roofline     -- the dominant kernel's achieved algorithmic bandwidth: a
                  breakdown pass outside the timed region brackets every
                  launch with HIP events on the library's stream
                  (emx_timing_begin/end) and names the kernel with the
                  largest total time; inside the timed region only that
                  kernel's launches carry events (emx_timing_select), and
                  it is reported with their average duration and its
                  algorithmic bytes per launch (each field it must read or
                  write once);
  cpu_baseline -- the reference's own Fortran (oracle/_ref, built from the
                  reference sources by flang -O3) running the same stage
                  once on one host core (it is serial code);
  parity       -- normwise relative error of the GPU stage outputs against
                  that Fortran run.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pts/s per FP64 RK stage + %HBM roofline; batched/learned vs plain at stated err"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", default="c6f")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-glue", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 256 MiB buffer 5 times (the reference kernel for "
                         "scaling hardware byte counters) and exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not "
              "supported (the multi-GPU sharding is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from emix3d_amd import Operator, fields

    if args.calibrate_copy:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": src.numel() * 8,
                          "bytes_written_per_launch": src.numel() * 8, "launches": 5}))
        return

    n = args.n
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    seed = 20261204
    t0 = time.time()
    f = fields.stage_fields(n, n, n, seed, curvilinear=True, turbulent=True, device=dev)
    op = Operator(n, n, n, args.scheme)
    jbn = op.jacobeano(f["x"])
    sgd, dmins, _ = op.inisgdm(f["x"])
    c4 = 1.0 / fields.REYNOLDS
    work = op.stage_buffers()
    torch.cuda.synchronize()
    log(f"bench: inputs ready in {time.time() - t0:.1f}s (n={n}, scheme={args.scheme})")

    def step():
        op.rk_stage(f["um"], f["u"], f["temp"], f["pres"], jbn, sgd, c4, work=work)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-kernel breakdown (events around every launch), outside the timed
    # region: the events add a gap per launch
    nb = max(3, min(args.steps, 10))
    op.timing_begin()
    for _ in range(nb):
        step()
    torch.cuda.synchronize()
    breakdown = kernel_breakdown(op.timing_end(), nb)
    top = dominant_label(breakdown)
    # timed region: events around the launches of the dominant kernel only
    op.timing_begin(label=top)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    live = [k for k in op.timing_end() if k["label"] == top]
    pts = float(n) ** 3
    value = pts / dt
    log(f"bench: {dt * 1e3:.3f} ms per stage, {value / 1e9:.3f} Gpts/s")
    roof = kernel_roofline(top, live, breakdown)
    for b in breakdown:
        log(f"bench:   {b['kernel']:44s} {b['us_per_launch']:8.1f} us x{b['launches_per_step']}"
            f"  {b['GBs']:6.0f} GB/s")

    # ---- stage algorithmic bytes (each distinct input read once, each output
    # written once): inputs um(5) u(3) temp pres jbn(1,5,9,10,11) sgd(7) = 22,
    # outputs dvel(9) dtemp(3) amut vis rs(5) rsv(5) = 24 fields of 8 B.
    stage_bytes_pt = (22 + 24) * 8
    stage_gbs = stage_bytes_pt * pts / dt / 1e9
    # the same stage counted per drop-in entry point with the flux
    # divergence fused per equation: gradient 128 (u, v, w, T in; 12 out),
    # sgdm + viscosidad 168, flujos + divergencia of the five equations 600
    fused_form_pt = 128 + 168 + 600
    fused_form_gbs = fused_form_pt * pts / dt / 1e9

    learned = None
    if not args.no_learned:
        learned = learned_section(op, f, jbn, sgd, c4, work, dt, args.steps, args.warmup)
    batched = None
    if not args.no_batched:
        batched = batched_section(args.steps, args.warmup)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_section(args.steps, args.warmup)
    small = None
    if not args.no_small:
        small = small_grid_section(args.steps, args.warmup)
    glue = None
    if not args.no_glue:
        glue = glue_section(op, f, jbn, sgd, c4, work, dt, args.steps, args.warmup)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(f, jbn, sgd, c4, work, n, args.scheme)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "curvilinear full RK-stage spatial operator (grad + sgdm "
                               "eddy viscosity + flux divergence of 5 equations)",
                   "grid": [n, n, n], "stretching": "tanh beta=2", "scheme": args.scheme,
                   "velocity": "solenoidal k^-5/3 (k<=32) + 6-blade radial jet",
                   "stage_algorithmic_bytes_per_pt": stage_bytes_pt,
                   "stage_algorithmic_GBs": stage_gbs,
                   "stage_frac_of_hbm_peak": stage_gbs / HBM_PEAK_GBS,
                   "stage_fused_form_bytes_per_pt": fused_form_pt,
                   "stage_fused_form_frac_of_hbm_peak": fused_form_gbs / HBM_PEAK_GBS},
        "roofline": roof,
        "kernels": breakdown,
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": batched,
        "learned": learned,
        "adjoint": adjoint,
        "small_grid": small,
        "substep_and_filter": glue,
    }
    print(json.dumps(out), flush=True)


# HBM bytes per launch of kernels measured with rocprofv3 --pmc (FETCH_SIZE
# and WRITE_SIZE, scaled by the same counters' reading for a copy kernel of
# known size, 1.9998 and 1.0), keyed by the kernel label (the accumulation
# passes: the average over their eta and zeta launches);
# profiles/rk_stage_256_c6f_hbm_traffic.md holds the runs.
MEASURED_TRAFFIC = {"flux_div_all[accumulate]": 881288263.0, "flux_project_xi": 6174124426.0,
                    "sgdm": 3716755246.0}


def kernel_breakdown(kernels, steps):
    """Per-kernel time and algorithmic bandwidth from the library's
    per-launch events, largest total time first."""
    breakdown = []
    for k in kernels:
        avg_ms = k["ms"] / k["launches"]
        b = k["bytes"] / k["launches"]
        name = k["label"] + (f" dir{k['dir']}" if k["dir"] >= 0 else "")
        breakdown.append({"kernel": name, "label": k["label"], "dir": k["dir"],
                          "launches_per_step": k["launches"] // steps,
                          "us_per_launch": avg_ms * 1e3, "bytes_per_launch": b,
                          "GBs": b / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0,
                          "ms_per_step": k["ms"] / steps})
    return breakdown


def dominant_label(breakdown):
    """The kernel label with the largest total time over all its launches
    and directions (the line kernels of one label share their code)."""
    tot = {}
    for b in breakdown:
        tot[b["label"]] = tot.get(b["label"], 0.0) + b["ms_per_step"]
    return max(tot, key=tot.get)


def kernel_roofline(label, live, breakdown):
    """Roofline of the dominant kernel: every launch of its label timed with
    HIP events on the library's stream inside the timed region (live)."""
    launches = sum(k["launches"] for k in live)
    ms = sum(k["ms"] for k in live)
    b = sum(k["bytes"] for k in live) / launches
    avg_ms = ms / launches
    gbs = b / (avg_ms * 1e-3) / 1e9
    share = (sum(x["ms_per_step"] for x in breakdown if x["label"] == label) /
             sum(x["ms_per_step"] for x in breakdown))
    return {"kernel": label, "bound": "hbm", "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get(label),
            "algorithmic_bytes_per_launch": b, "avg_launch_ms": avg_ms,
            "launches_timed": launches, "share_of_step": share}


def timed(fn, steps, warmup):
    """Seconds per call of fn (whole-GPU work), synchronised on both sides."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def rel_l2(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def glue_section(op, f, jbn, sgd, c4, work, stage_s, steps, warmup):
    """One whole device RUNGEK sub-step (operator + SSP-RK3 update +
    boundary values + velocity clip + pressure extrapolation into a solid +
    rebuild) and the compact filter FILTRADO, at the benchmark size."""
    import torch
    from emix3d_amd import fields
    n = f["u"].shape[-1]
    N = float(n) ** 3
    mach, dt = 0.3, 1e-3
    gen = torch.Generator(device="cuda").manual_seed(3)
    st = {k: f[k].clone() for k in ("um", "u", "temp", "pres")}
    um0 = f["um"].clone()
    mask = (torch.rand(f["pres"].shape, dtype=torch.float64, device="cuda",
                       generator=gen) > 0.02).double()
    # a solid ball of radius 0.15 at the centre of the unit box (sobject > 0
    # inside, the signed distance outside): about 1.4% of the points
    x = fields.grid_uniform(n, n, n, device=torch.device("cuda"))
    sobject = (0.15 - torch.sqrt(((x - 0.5) ** 2).sum(0))).contiguous()
    dxyz = (1.0 / (n - 1),) * 3
    um1 = op._empty(5)
    work = op.stage_buffers()  # its own: the stage's buffers are checked for parity later

    def substep():
        op.rk_substep(2, um0, st["um"], st["u"], st["temp"], st["pres"], jbn, sgd, c4, dt, mach,
                      mask, sobject=sobject, dxyz=dxyz, deltamin=dxyz[0], work=work, um1=um1)

    t_sub = timed(substep, steps, warmup)
    op.timing_begin()
    substep()
    torch.cuda.synchronize()
    glue_k = {}
    for k in op.timing_end():
        if k["label"] in ("rk_update", "frontera_z", "frontera_y", "frontera_x", "rk_clip",
                          "rk_rebuild") or k["label"].startswith("ls_"):
            key = k["label"] + (f" dir{k['dir']}" if k["dir"] >= 0 else "")
            glue_k[key] = {"launches": k["launches"], "us": k["ms"] * 1e3,
                           "GBs": k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0}
    fu = {"um1": f["um"].clone()}
    u3, p1, t1 = op._empty(3), op._empty(), op._empty()

    def filt():
        op.filtrado(1, fu["um1"], jbn, 0.49, mach, u=u3, pres=p1, temp=t1)

    t_f = timed(filt, max(3, steps // 2), 1)
    # level-set reinitialisation of a distorted distance to a sphere
    r = torch.sqrt(((x - 0.5) ** 2).sum(0))
    phi_init = ((r - 0.3) * (1.0 + 2.0 * (r - 0.3))).contiguous()
    phi = phi_init.clone()
    dmin = torch.full_like(phi, dxyz[0])

    def reinit():
        phi.copy_(phi_init)
        op.reinit_level_set(phi, jbn, dmin, dxyz, dxyz[0])

    t_r = timed(reinit, max(3, steps // 4), 1)
    NG = float(n + 10) ** 3
    # per iteration: ghost mirror (16 B) and three stages (56 + 64 + 64 B) on
    # the padded grid; five paddings and the unpadding; the copy of phi_init
    r_bytes = (10 * (16 + 184) + 5 * 16) * NG + 16 * N + 16 * N
    # FILTRADO: 15 line passes (field, J, 1/J in; field out) + the
    # primitive variables (um1 1..5 in; u, pres, temp out)
    f_bytes = (15 * 4 + 10) * 8.0 * N
    out = {"workload": f"{n}^3 device RUNGEK sub-step 2 (operator, update, boundary values, "
                       "clip, pressure extrapolation into a central solid ball of radius 0.15, "
                       "rebuild), FILTRADO(itf=1, cfilt=0.49) and the level-set "
                       "reinitialisation (10 iterations) of a distorted sphere distance",
           "substep_ms": t_sub * 1e3, "stage_ms": stage_s * 1e3,
           "substep_pts_per_s": N / t_sub, "glue_kernels": glue_k,
           "filter_ms": t_f * 1e3, "filter_GBs_algorithmic": f_bytes / t_f / 1e9,
           "filter_frac_of_hbm_peak": f_bytes / t_f / 1e9 / HBM_PEAK_GBS,
           "reinit_ms": t_r * 1e3, "reinit_GBs_algorithmic": r_bytes / t_r / 1e9}
    log(f"bench: sub-step and filter {out}")
    return out


def learned_section(op, f, jbn, sgd, c4, work, plain_s, steps, warmup):
    """The 256^3 stage with the eddy viscosity from the learned closure
    (BF16 matrix cores) against the same stage with dynamic Smagorinsky
    computed directly (the model the network learns) and against the plain
    stage (structure-function sgdm, the reference's model, timed above):
    throughput of each, the closure kernels alone, and the
    relative L2 error of nu_t and of the viscous flux divergence rsv
    (equations 2..4) against the dynamic model, for the network, the static
    Smagorinsky model (one least-squares constant fitted to the training
    targets) and sgdm."""
    import torch
    from emix3d_amd.closure import WEIGHTS, LearnedClosure, StaticSmagorinsky
    if not os.path.exists(WEIGHTS):
        log("bench: no trained closure weights; learned section skipped")
        return None
    cl = LearnedClosure.load()
    delta = sgd[0].contiguous()
    u, um, temp, pres = f["u"], f["um"], f["temp"], f["pres"]
    # own buffers: the stage's outputs stay as the plain stage left them
    dv, dt_ = torch.empty_like(work["dvel"]), torch.empty_like(work["dtemp"])
    vis, rs, rsv = (torch.empty_like(work["vis"]), torch.empty_like(work["rs"]),
                    torch.empty_like(work["rsv"]))
    nut_d = torch.empty_like(delta)
    nut_m = torch.empty_like(delta)

    def stage(closure):
        op.deriv_vel(u, temp, dvel=dv, dtemp=dt_)
        if closure == "dyn":
            op.dynsmag(u, dv, jbn, delta, nut=nut_d, c4=c4, vis=vis)
        else:
            cl.nut(op, dv, jbn, delta, out=nut_m, c4=c4, vis=vis)
        op.flujos_divergencia_all(um, u, pres, vis, jbn, dv, rs=rs, rsv=rsv)

    t_dyn = timed(lambda: stage("dyn"), steps, warmup)
    rsv_dyn = rsv[1:4].clone()
    t_mlp = timed(lambda: stage("mlp"), steps, warmup)
    rsv_mlp = rsv[1:4].clone()
    # the closures alone
    op.timing_begin()
    cl.nut(op, dv, jbn, delta, out=nut_m, c4=c4, vis=vis)
    op.dynsmag(u, dv, jbn, delta, nut=nut_d, c4=c4, vis=vis)
    torch.cuda.synchronize()
    kern = {k["label"]: {"us": k["ms"] * 1e3, "GBs": k["bytes"] / (k["ms"] * 1e-3) / 1e9}
            for k in op.timing_end() if k["label"] in ("mlp_infer", "dynsmag_lm", "dynsmag_nut")}
    pts = float(delta.numel())
    if "mlp_infer" in kern:
        kern["mlp_infer"]["TFLOPs"] = 17792.0 * pts / (kern["mlp_infer"]["us"] * 1e-6) / 1e12
    amut, _, _ = op.viscosidad_t(u, dv, jbn, sgd, c4, want_vis5=False)
    op.flujos_divergencia_all(um, u, pres, (c4 + amut).contiguous(), jbn, dv, rs=rs, rsv=rsv)
    rsv_sgdm = rsv[1:4].clone()
    out = {"workload": "256^3 curvilinear stage (grad + closure + flux divergence of 5 eqs)",
           "plain_sgdm_pts_per_s": pts / plain_s,
           "dynamic_smagorinsky_pts_per_s": pts / t_dyn,
           "learned_pts_per_s": pts / t_mlp,
           "learned_vs_plain": plain_s / t_mlp,
           "learned_vs_dynamic": t_dyn / t_mlp,
           "closure_kernels": kern,
           "nut_rel_l2_vs_dynamic": rel_l2(nut_m, nut_d),
           "rsv_rel_l2_vs_dynamic": rel_l2(rsv_mlp, rsv_dyn),
           "sgdm_nut_rel_l2_vs_dynamic": rel_l2(amut, nut_d),
           "sgdm_rsv_rel_l2_vs_dynamic": rel_l2(rsv_sgdm, rsv_dyn),
           "inference": "BF16 MFMA (16x16x32), FP32 accumulation, FP64 stage"}
    if cl.static_c is not None:
        nut_s = StaticSmagorinsky(cl.static_c).nut(dv, jbn, delta)
        op.flujos_divergencia_all(um, u, pres, (c4 + nut_s).contiguous(), jbn, dv, rs=rs, rsv=rsv)
        out["static_smagorinsky_c"] = cl.static_c
        out["static_nut_rel_l2_vs_dynamic"] = rel_l2(nut_s, nut_d)
        out["static_rsv_rel_l2_vs_dynamic"] = rel_l2(rsv[1:4], rsv_dyn)
    log(f"bench: learned closure {out}")
    return out


def adjoint_section(steps, warmup):
    """The adjoint of the composed operator (velocity gradient, fluxes of
    the five equations, divergence) on a 128^3 tanh-stretched grid with the
    sine-mode field: emx_flux_div_all_adj + emx_grad_adj of random
    cotangents, points per second and the library's algorithmic bytes per
    launch."""
    import torch
    from emix3d_amd import Operator, fields
    n = 128
    f = fields.stage_fields(n, n, n, 20261206, curvilinear=True, turbulent=False, device="cuda")
    op = Operator(n, n, n, "c6c")
    jbn = op.jacobeano(f["x"])
    gen = torch.Generator(device="cuda").manual_seed(4)
    vis = 1e-3 + 1e-3 * torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rb = torch.randn((5, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    rvb = torch.randn((5, n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    dvel, _ = op.deriv_vel(f["u"])
    outb = None

    def adj():
        nonlocal outb
        outb = op.flujos_divergencia_all_adj(rb, rvb, f["um"], f["u"], vis, jbn, dvel, out=outb)
        op.deriv_vel_adj(outb["dvel_bar"])

    t = timed(adj, steps, warmup)
    op.timing_begin()
    adj()
    torch.cuda.synchronize()
    ks = op.timing_end()
    ms = sum(k["ms"] for k in ks)
    byt = sum(k["bytes"] for k in ks)
    pts = float(n) ** 3
    out = {"workload": "128^3 tanh-stretched adjoint of grad + fluxes + divergence (5 eqs)",
           "pts_per_s": pts / t, "ms": t * 1e3,
           "algorithmic_bytes_per_pt": byt / pts,
           "GBs_algorithmic": byt / (ms * 1e-3) / 1e9,
           "frac_of_hbm_peak": byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernels": {k["label"] + (f" dir{k['dir']}" if k["dir"] >= 0 else ""):
                       {"launches": k["launches"], "us": k["ms"] * 1e3,
                        "GBs": k["bytes"] / (k["ms"] * 1e-3) / 1e9} for k in ks}}
    log(f"bench: adjoint {out}")
    return out


def small_grid_section(steps, warmup):
    """The 64^3 uniform grid with 8 sine modes per axis: the C6F velocity
    gradient fused with the velocity divergence (emx_grad_div).  One field
    is 2 MiB, so the working set sits in the 256 MB last-level cache: points
    per second only, no HBM roofline."""
    import torch
    from emix3d_amd import Operator, fields
    n = 64
    x = fields.grid_uniform(n, n, n, device="cuda")
    u = fields.velocity_sines(x, 20261203).contiguous()
    op = Operator(n, n, n, "c6f")
    jbn = op.jacobeano(x)
    dvel, div = op.grad_div(u, jbn=jbn)
    t = timed(lambda: op.grad_div(u, jbn=jbn, dvel=dvel, div=div), max(steps, 50), warmup)
    out = {"workload": "64^3 uniform, 8 sine modes per axis: C6F gradient + velocity divergence",
           "pts_per_s": float(n) ** 3 / t, "us": t * 1e6}
    log(f"bench: small grid {out}")
    return out


def batched_section(steps, warmup):
    """48 boxes of uneven sizes (32^3 .. 192x128x96), each with the sine-mode
    velocity and a sphere-plus-free-surface level set: one stage = the
    gradient fused with the velocity divergence plus one WENO5-RK3 level-set
    step (CFL 0.5), for all boxes.  Batched (one launch per phase, HIP graph,
    tuned line shapes) against the plain per-box form (the single-grid entry
    points box after box), with the largest difference between the two."""
    import torch
    from emix3d_amd import Batch, Operator, fields
    dims = fields.batch_dims(20261205, 48)
    boxes, ops = [], []
    b = Batch(dims, "c6c")
    for i, d in enumerate(dims):
        z = fields.batch_box(d, 1000 + i, device="cuda")
        op = Operator(*d, "c6c")
        z["jbn"] = op.jacobeano(z["x"])
        del z["x"]
        nx, ny, nz = d
        z["phi0"] = z["phi"].clone()
        for k in ("dvel", "dvel_p"):
            z[k] = torch.empty((9, nz, ny, nx), dtype=torch.float64, device="cuda")
        for k in ("div", "div_p"):
            z[k] = torch.empty((nz, ny, nx), dtype=torch.float64, device="cuda")
        z["phi_p"] = z["phi"].clone()
        b.bind(i, z["u"], z["jbn"], z["phi"], z["dvel"], z["div"], z["ih"], z["dxyz"], z["dt"])
        boxes.append(z)
        ops.append(op)
    pts = float(sum(nx * ny * nz for nx, ny, nz in dims))

    def plain():
        for z, op in zip(boxes, ops):
            op.grad_div(z["u"], ih=z["ih"], jbn=z["jbn"], dvel=z["dvel_p"], div=z["div_p"])
            op.advect_weno5(z["phi_p"], z["u"], z["jbn"], z["dxyz"], z["dt"])

    b.tune(reps=3)
    plans = [b.get_plan(d) for d in range(3)]
    for z in boxes:
        z["phi"].copy_(z["phi0"])
        z["phi_p"].copy_(z["phi0"])
    b.stage(graph=True)
    plain()
    torch.cuda.synchronize()
    err = 0.0
    for z in boxes:
        for a, p in (("dvel", "dvel_p"), ("div", "div_p"), ("phi", "phi_p")):
            err = max(err, float((z[a] - z[p]).abs().max() / z[p].abs().max()))
    t_plain = timed(plain, steps, warmup)
    t_batch = timed(lambda: b.stage(graph=True), steps, warmup)
    t_eager = timed(lambda: b.stage(graph=False), steps, warmup)
    bytes_pt = 240.0  # gradient 96 + divergence 72 + WENO5 step 72 B per point
    out = {"workload": "48 boxes 32^3..192x128x96: grad + velocity divergence + WENO5-RK3 "
                       "level-set step", "boxes": len(dims), "points": pts,
           "plain_per_box_pts_per_s": pts / t_plain,
           "batched_graph_pts_per_s": pts / t_batch,
           "batched_eager_pts_per_s": pts / t_eager,
           "speedup_vs_plain": t_plain / t_batch,
           "batched_GBs_algorithmic": bytes_pt * pts / t_batch / 1e9,
           "frac_of_hbm_peak": bytes_pt * pts / t_batch / 1e9 / HBM_PEAK_GBS,
           "max_normwise_diff_vs_plain": err,
           "line_shapes_lg_C": plans}
    log(f"bench: batched {out}")
    return out


def cpu_baseline(f, jbn, sgd, c4, work, n, scheme):
    """The reference Fortran stage (oracle/_ref) once, on one core."""
    import numpy as np
    import torch
    try:
        from oracle.pyoracle import RefFortran
    except Exception as e:  # pragma: no cover
        log(f"bench: cpu baseline unavailable ({e})")
        return None, None
    lib_scheme = scheme if scheme in ("c4", "c6f") else None
    if lib_scheme is None or not RefFortran.available(lib_scheme):
        log("bench: reference Fortran build not present; no cpu baseline")
        return None, None
    h = {k: np.ascontiguousarray(v.cpu().numpy()) for k, v in f.items()}
    hj = np.ascontiguousarray(jbn.cpu().numpy())
    hs = np.ascontiguousarray(sgd.cpu().numpy())
    ref = RefFortran(lib_scheme)
    ref.init(n, n, n)
    # INISGDM sets the model constants (isgm, amu0, cutoff) of the run;
    # its geometry fields must equal the GPU's (checked here)
    rsgd, _, _ = ref.inisgdm(h["x"])
    sgd_err = float(np.abs(rsgd - hs).max() / np.abs(rsgd).max())
    ref.stage_load(h["um"], h["u"], h["temp"], h["pres"], hj, hs, c4)
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    log("bench: timing the reference Fortran stage on one core ...")
    sys.stdout.flush()
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)  # the reference prints per call
    try:
        t0 = time.perf_counter()
        ref.stage_run()
        t = time.perf_counter() - t0
        ref.flush()
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    rs, rsv, amut, vis, dvel = ref.stage_fetch()
    log(f"bench: reference stage {t:.2f}s on 1 core")

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    g = {k: v.cpu().numpy() for k, v in work.items()}
    parity = {"vs": "reference Fortran stage (same inputs, scheme %s)" % lib_scheme,
              "sgd_setup": sgd_err,
              "dvel": rel(g["dvel"], dvel), "amut": rel(g["amut"], amut),
              "vis": rel(g["vis"], vis)}
    for m in range(5):
        parity[f"rs{m + 1}"] = rel(g["rs"][m], rs[m])
        if m in (1, 2, 3):
            parity[f"rsv{m + 1}"] = rel(g["rsv"][m], rsv[m])
    parity["max"] = max(v for k, v in parity.items() if k != "vs")
    cpu = {"value": float(n) ** 3 / t, "unit": "grid-pts/s", "cores": 1, "kind": "reference",
           "sample": f"one full {n}^3 RK-stage operator of the reference Fortran "
                     "(deriv_vel, viscosidad_t, viscosidad, 5x flujos+divergencia; "
                     "flang -O3, serial)", "seconds": t}
    return cpu, parity


if __name__ == "__main__":
    main()
