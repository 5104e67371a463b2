"""Learned eddy-viscosity closure: training data, training loop, inference.

The network (10 -> 64 -> 64 -> 64 -> 1, SiLU hidden activations, softplus
output) maps the physical velocity gradient du_a/dx_b and the filter width
to nu_t >= 0.  The reference has no such closure; its model is the
structure-function sgdm (solver/sgdm_incomp.f90), which stays the fallback
and the judge of the network's error.

Training runs on the GPU through the library's own kernels:
  * data: snapshots of the sine-mode velocity family (8 modes per axis,
    seeded amplitudes and phases) on a tanh-stretched grid; the target is
    the dynamic Smagorinsky nu_t (emx_dynsmag, box test filter of ratio 2);
  * loss: the normalised squared error of nu_t at sampled points plus the
    normalised squared error of the viscous flux divergence rsv of the
    five-equation operator with vis = 1/Re + nu_t.  The gradient of the
    second term with respect to nu_t is the adjoint kernel's vis_bar
    (emx_flux_div_all_adj);
  * network gradients: emx_mlp_forward_f32 / emx_mlp_backward_f32 (FP32),
    parameter update emx_adam_f32.
Inference (emx_mlp_nut) runs the hidden layers on the matrix cores with BF16
operands and FP32 accumulation.
"""
import math
import os
import time

import numpy as np
import torch

from . import fields
from .api import (EMX_MLP_NNORM, EMX_MLP_NPARAM, Operator, adam_f32, mlp_backward_f32,
                  mlp_forward_f32)

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_weights.npz")
C4 = 1.0 / fields.REYNOLDS


def init_params(seed):
    """Uniform Glorot initialisation, zero biases; the output bias puts
    softplus at 1 (nu = out_scale) for zero input."""
    rng = np.random.default_rng(seed)
    p = np.zeros(EMX_MLP_NPARAM, dtype=np.float32)
    off = 0
    for fan_in, fan_out in ((10, 64), (64, 64), (64, 64), (64, 1)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        p[off:off + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        off += fan_in * fan_out + fan_out
    p[-1] = math.log(math.e - 1.0)
    return p


class SnapshotFamily:
    """Velocity snapshots on one tanh-stretched grid, with the grid's
    metrics, filter width (the sgdm width dxyzsgd) and the
    structure-function model's geometry.  kind "sines": the sine-mode family
    (8 modes per axis, seeded amplitudes and phases); "turbulent": the
    solenoidal k^-5/3 field with the impeller jet.  seeds: the snapshot seeds
    training cycles through; cache keeps generated snapshots on the GPU."""

    def __init__(self, n, scheme="c6c", device="cuda", kind="sines", seeds=range(1, 201),
                 cache=False):
        self.n = n
        self.kind = kind
        self.seeds = list(seeds)
        self.cache = {} if cache else None
        self.device = torch.device(device)
        self.op = Operator(n, n, n, scheme)
        self.x = fields.grid_tanh(n, n, n, device=self.device)
        self.jbn = self.op.jacobeano(self.x)
        self.sgd, _, _ = self.op.inisgdm(self.x)
        self.delta = self.sgd[0].contiguous()
        zero = torch.zeros((n, n, n), dtype=torch.float64, device=self.device)
        self.um0 = torch.zeros((5, n, n, n), dtype=torch.float64, device=self.device)
        self.pres0 = zero

    def snapshot(self, seed):
        if self.cache is not None and seed in self.cache:
            return self.cache[seed]
        if self.kind == "sines":
            u = fields.velocity_sines(self.x, seed).contiguous()
        else:
            u = fields.velocity_turbulent(self.x, seed).contiguous()
        dvel, _ = self.op.deriv_vel(u)
        nut = self.op.dynsmag(u, dvel, self.jbn, self.delta)
        if self.cache is not None:
            self.cache[seed] = (u, dvel, nut)
        return u, dvel, nut

    def rsv(self, u, dvel, nut):
        """Viscous flux divergence of the five equations (2..4 non-zero)."""
        vis = (C4 + nut).contiguous()
        _, rsv = self.op.flujos_divergencia_all(self.um0, u, self.pres0, vis, self.jbn, dvel)
        return rsv


class LearnedClosure:
    def __init__(self, params=None, norm=None, device="cuda"):
        self.device = torch.device(device)
        self.params = None if params is None else torch.as_tensor(
            np.asarray(params, np.float32), device=self.device).contiguous()
        self.norm = None if norm is None else torch.as_tensor(
            np.asarray(norm, np.float32), device=self.device).contiguous()
        self.static_c = None

    @classmethod
    def load(cls, path=WEIGHTS, device="cuda"):
        """The network, and the static Smagorinsky constant fitted to the
        same training targets (static_c, None in older files)."""
        z = np.load(path, allow_pickle=False)
        cl = cls(z["params"], z["norm"], device)
        cl.static_c = float(z["static_c"]) if "static_c" in z.files else None
        return cl

    def save(self, path=WEIGHTS, **meta):
        np.savez(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy(),
                 **{k: np.asarray(v) for k, v in meta.items()})

    def nut(self, op, dvel, jbn, delta, out=None, c4=0.0, vis=None):
        return op.mlp_nut(self.params, self.norm, dvel, jbn, delta, nut=out, c4=c4, vis=vis)

    # ------------------------------------------------------------------ fit
    def fit_normalisation(self, fams, per_family=4, samples=65536, gen=None):
        feats, tg = [], []
        for fam in fams:
            N = fam.n ** 3
            for s in fam.seeds[:per_family]:
                u, dvel, nut = fam.snapshot(s)
                idx = torch.randint(0, N, (samples,), device=self.device, generator=gen)
                feats.append(fam.op.mlp_gather(dvel, fam.jbn, fam.delta, idx))
                tg.append(nut.reshape(-1)[idx])
        f = torch.cat(feats).double()
        mean = f.mean(0)
        std = f.std(0).clamp_min(1e-30)
        out_scale = torch.cat(tg).mean().clamp_min(1e-30)
        norm = torch.cat([mean, 1.0 / std, out_scale.reshape(1)]).float()
        assert norm.numel() == EMX_MLP_NNORM
        self.norm = norm.contiguous()

    def train(self, fams, steps=600, batch=65536, lr=2e-3, flux_weight=0.5, seed=20261208,
              log=None):
        """Adam on the two-term loss; the families take turns, each cycling
        through its seeds."""
        if isinstance(fams, SnapshotFamily):
            fams = [fams]
        gen = torch.Generator(device=self.device).manual_seed(seed)
        if self.params is None:
            self.params = torch.as_tensor(init_params(seed), device=self.device).contiguous()
        if self.norm is None:
            self.fit_normalisation(fams, gen=gen)
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        work = None
        hist = []
        t0 = time.time()
        for step in range(steps):
            fam = fams[step % len(fams)]
            k = step // len(fams)
            u, dvel, target = fam.snapshot(fam.seeds[k % len(fam.seeds)])
            N = fam.n ** 3
            idx = torch.randint(0, N, (batch,), device=self.device, generator=gen)
            feat = fam.op.mlp_gather(dvel, fam.jbn, fam.delta, idx)
            act, nu = mlp_forward_f32(self.params, self.norm, feat)
            tgt = target.reshape(-1)[idx].float()
            den = float((target * target).mean())
            diff = nu - tgt
            dnu = (2.0 / (batch * den)) * diff
            loss_nu = float((diff * diff).mean()) / den
            loss_f = 0.0
            if flux_weight > 0.0:
                nut_full = self.nut(fam.op, dvel, fam.jbn, fam.delta)
                rsv = fam.rsv(u, dvel, nut_full)
                rsv_t = fam.rsv(u, dvel, target)
                d = rsv[1:4] - rsv_t[1:4]
                den_f = float((rsv_t[1:4] * rsv_t[1:4]).sum())
                loss_f = float((d * d).sum()) / den_f
                rsv_bar = torch.zeros_like(rsv)
                rsv_bar[1:4] = (2.0 / den_f) * d
                vis = (C4 + nut_full).contiguous()
                bars = fam.op.flujos_divergencia_all_adj(torch.zeros_like(rsv), rsv_bar,
                                                         fam.um0, u, vis, fam.jbn, dvel)
                dnu = dnu + (flux_weight * N / batch) * bars["vis_bar"].reshape(-1)[idx].float()
            grad, work = mlp_backward_f32(self.params, self.norm, feat, act, dnu.contiguous(),
                                          work)
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / steps))
            adam_f32(self.params, grad, m, v, step + 1, lr=lr_t)
            hist.append((loss_nu, loss_f))
            if log and (step % 50 == 0 or step == steps - 1):
                log(f"closure train step {step}: loss nu {loss_nu:.4e} flux {loss_f:.4e} "
                    f"({time.time() - t0:.1f}s)")
        return hist


def rel_l2(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def strain_scale(dvel, jbn, delta):
    """delta^2 |S| with |S| = sqrt(2 S_ij S_ij) of the physical gradient
    du_a/dx_b = jbn(1+4b) dvel(3b+a): the static Smagorinsky model is
    nu_t = C * delta^2 |S|."""
    m = (jbn[0], jbn[4], jbn[8])
    s2 = torch.zeros_like(delta)
    for a in range(3):
        for b in range(3):
            sab = 0.5 * (m[b] * dvel[3 * b + a] + m[a] * dvel[3 * a + b])
            s2 += sab * sab
    return delta * delta * torch.sqrt(2.0 * s2)


class StaticSmagorinsky:
    """nu_t = C delta^2 |S| with one constant C, fitted by least squares to
    the dynamic-Smagorinsky targets of the training snapshots: the baseline
    a network fed the whole velocity gradient has to beat."""

    def __init__(self, c=None):
        self.c = c

    def fit(self, fams, per_family=8):
        num = den = 0.0
        for fam in fams:
            for s in fam.seeds[:per_family]:
                _, dvel, nut = fam.snapshot(s)
                q = strain_scale(dvel, fam.jbn, fam.delta)
                num += float((q * nut).sum())
                den += float((q * q).sum())
        self.c = num / den
        return self

    def nut(self, dvel, jbn, delta):
        return self.c * strain_scale(dvel, jbn, delta)


def evaluate(fam, closure, seeds, static=None):
    """Relative L2 errors against dynamic Smagorinsky on the snapshots
    seeds: of nu_t and of the viscous flux divergence rsv (equations 2..4),
    for the learned closure, for the static Smagorinsky baseline (when
    given) and for the structure-function sgdm."""
    names = ["mlp", "sgdm"] + (["static"] if static is not None else [])
    out = {f"{n}_{q}": [] for n in names for q in ("nut", "rsv")}
    for s in seeds:
        u, dvel, target = fam.snapshot(s)
        rsv_t = fam.rsv(u, dvel, target)[1:4]
        models = {"mlp": closure.nut(fam.op, dvel, fam.jbn, fam.delta)}
        models["sgdm"], _, _ = fam.op.viscosidad_t(u, dvel, fam.jbn, fam.sgd, C4, want_vis5=False)
        if static is not None:
            models["static"] = static.nut(dvel, fam.jbn, fam.delta)
        for name, nu in models.items():
            out[f"{name}_nut"].append(rel_l2(nu, target))
            out[f"{name}_rsv"].append(rel_l2(fam.rsv(u, dvel, nu)[1:4], rsv_t))
    return {k: float(np.mean(v)) for k, v in out.items()}


def training_families(device="cuda"):
    """The training data: 200 snapshots of the sine-mode family at 128^3
    (seeds 1..200), and the turbulent family the RK stage runs on, at 128^3
    and 256^3 (seeds 1001.., kept on the GPU once generated), so that the
    network sees the gradient statistics and filter widths it is used at."""
    return [SnapshotFamily(128, kind="sines", seeds=range(1, 201), device=device),
            SnapshotFamily(128, kind="turbulent", seeds=range(1001, 1017), cache=True,
                           device=device),
            SnapshotFamily(256, kind="turbulent", seeds=range(2001, 2007), cache=True,
                           device=device)]


def main():
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="train the learned closure on the GPU")
    ap.add_argument("--steps", type=int, default=1200)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    fams = training_families()
    cl = LearnedClosure()
    hist = cl.train(fams, steps=args.steps, batch=args.batch, log=log)
    static = StaticSmagorinsky().fit(fams)
    log(f"static Smagorinsky constant (least squares on the training targets): {static.c:.5f}")
    evals = {}
    held = {"sines128": (fams[0], range(201, 205)),
            "turbulent128": (SnapshotFamily(128, kind="turbulent"), range(1101, 1103)),
            "turbulent256": (SnapshotFamily(256, kind="turbulent"), range(2101, 2103))}
    for name, (fam, seeds) in held.items():
        ev = evaluate(fam, cl, seeds, static)
        log(f"held-out {name} {list(seeds)}: {ev}")
        for k, v in ev.items():
            evals[f"{name}_{k}"] = v
    cl.save(args.out, steps=args.steps, loss=np.asarray(hist), static_c=static.c,
            eval_keys=np.asarray(list(evals.keys())), eval_vals=np.asarray(list(evals.values())))
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
