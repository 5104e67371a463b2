"""torch.autograd wrappers of the operator, for the verification harness
only: forward through the C ABI's forward entry points, backward through
its adjoint entry points (emx_grad_adj, emx_flux_div_all_adj).  The library
itself exports plain C functions; nothing in emix3d_amd depends on torch
autograd."""
import torch


class FluxDivOfState(torch.autograd.Function):
    """(rs, rsv) = five-equation flux divergence of (um, u, pres, vis), with
    dvel = DERIV_VEL(u) inside; jbn fixed."""

    @staticmethod
    def forward(ctx, op, jbn, um, u, pres, vis):
        dvel, _ = op.deriv_vel(u)
        rs, rsv = op.flujos_divergencia_all(um, u, pres, vis, jbn, dvel)
        ctx.op = op
        ctx.save_for_backward(jbn, um, u, vis, dvel)
        return rs, rsv

    @staticmethod
    def backward(ctx, rs_bar, rsv_bar):
        op = ctx.op
        jbn, um, u, vis, dvel = ctx.saved_tensors
        b = op.flujos_divergencia_all_adj(rs_bar.contiguous(), rsv_bar.contiguous(), um, u, vis,
                                          jbn, dvel)
        ub, _ = op.deriv_vel_adj(b["dvel_bar"])
        return None, None, b["um_bar"], b["u_bar"] + ub, b["pres_bar"], b["vis_bar"]


class Derivative(torch.autograd.Function):
    """Compact first derivative along one direction."""

    @staticmethod
    def forward(ctx, op, f, direction):
        ctx.op = op
        ctx.direction = direction
        return op.derivative(f, direction)

    @staticmethod
    def backward(ctx, g):
        return None, ctx.op.derivative_adj(g.contiguous(), ctx.direction), None
