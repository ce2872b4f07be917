"""Differentiable forward NUFFT for torch.

``NufftForward.apply(x, om, A)`` returns ``y = A(om) x`` (``A.forward``) with
hand-written gradients:

* image: ``dL/dx = A^H g = prod(Kd) * A.adjoint(g)`` (the reference's adjoint
  divides by prod(Kd), ``nufft/_nufft_class_methods_cpu.py:181-193``);
* trajectory: ``dL/dom[m, d] = Re(conj(g[m]) dy[m]/dom[m, d])`` from the
  device kernel behind ``A.forward_grad_om`` (product rule over the per-axis
  min-max factors).

``om`` must be the coordinates ``A`` was planned with (the plan is a
function of om); to take a step in om, re-plan.  ``g`` follows torch's
complex convention (dL/d conj(y)), so a real loss gives real gradients.
"""
import numpy as np
import torch


class NufftForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, om, A):
        xt = x.detach().to(device=A.device, dtype=torch.complex64).contiguous()
        k = A.xx2k(A.x2xx(xt))
        y = A.k2y(k)
        ctx.A = A
        ctx.om_dtype = om.dtype if isinstance(om, torch.Tensor) else None
        ctx.save_for_backward(k)
        return y

    @staticmethod
    def backward(ctx, g):
        A = ctx.A
        (k,) = ctx.saved_tensors
        g = g.to(torch.complex64).contiguous()
        gx = gom = None
        if ctx.needs_input_grad[0]:
            gx = A.adjoint(g) * float(np.prod(A.Kd))
        if ctx.needs_input_grad[1]:
            gom = A.forward_grad_om(k, g)
            if ctx.om_dtype is not None:
                gom = gom.to(ctx.om_dtype)
        return gx, gom, None


def nufft_forward(A, x, om=None):
    """y = A(om) x, differentiable in x and (if given as a tensor) om."""
    if om is None:
        om = torch.zeros(0)
    return NufftForward.apply(x, om, A)
