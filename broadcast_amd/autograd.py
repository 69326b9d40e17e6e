"""The DNC5 residual as a differentiable torch operator.

``residual(w, flow)`` returns R(w) (interior rows; ghost rows are 0) for a
flat float64 HIP tensor ``w`` in the reference's layout, ghost frame
included.  As in the reference's tangent and adjoint routines the ghost
frame is an independent input (boundary conditions are applied by the
caller), so

* reverse mode calls the hand-written VJP kernels (bca_vjp_dnc5, replacing
  srcfv/adjoint/flux_num_dnc5_b.f90:15): grad_w = (dR/dw)^T grad_R;
* forward mode calls the tangent kernels (bca_jvp_dnc5, replacing
  srcfv/tangent/flux_num_dnc5_d.f90:15): R_dot = dR/dw . w_dot.

Both are exact derivatives of the same arithmetic, so gradcheck-style tests
compare them with finite differences, with each other (the dot-product
test) and with the reference's generated routines.
"""
from __future__ import annotations

from . import device as dev

torch = dev.require_gpu()


class Residual(torch.autograd.Function):
    @staticmethod
    def forward(w, flow):
        return flow.residual(w.detach())

    @staticmethod
    def setup_context(ctx, inputs, output):
        w, flow = inputs
        ctx.flow = flow
        ctx.save_for_backward(w)
        ctx.save_for_forward(w)

    @staticmethod
    def backward(ctx, grad_res):
        (w,) = ctx.saved_tensors
        wb = ctx.flow.vjp(w.detach(), grad_res.detach().contiguous())
        return wb, None

    @staticmethod
    def jvp(ctx, w_dot, _flow_dot):
        (w,) = ctx.saved_tensors
        _, resd = ctx.flow.jvp_flux(w.detach(), w_dot.detach().contiguous())
        return resd


def residual(w, flow):
    """R(w) with reverse- and forward-mode derivatives through the HIP kernels."""
    return Residual.apply(w, flow)
