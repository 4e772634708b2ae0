"""torch.autograd.Function over the GPU evaluator: phi = FMM(q, y; x) with a hand-written backward that
runs the adjoint FMM pass (exafmm.adjoint).

Tensors are CPU float64: the evaluator runs on the GPU through its own HIP runtime, and a process that
also initialises torch's GPU runtime (any torch.cuda call) leaves the device to that runtime, so the data
crosses as host arrays.  Real kernels only (Laplace, modified Helmholtz).
"""
import torch

import exafmm
from exafmm import adjoint as _adj


class FMMPotential(torch.autograd.Function):
    """phi (ntrg,) at targets x (ntrg, 3) of charges q (nsrc,) at sources y (nsrc, 3).

    forward(q, y, x, kernel, p, ncrit, wavek, precision); gradients for q and y (x is treated as data)."""

    @staticmethod
    def forward(ctx, q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
        qn = q.detach().cpu().double().numpy()
        yn = y.detach().cpu().double().numpy()
        xn = x.detach().cpu().double().numpy()
        v = _adj.evaluate(kernel, yn, qn, xn, p, ncrit, wavek, precision)
        ctx.save_for_backward(q, y, x)
        ctx.args = (kernel, p, ncrit, wavek, precision)
        return torch.from_numpy(v[:, 0].copy()).to(q.dtype)

    @staticmethod
    def backward(ctx, gphi):
        q, y, x = ctx.saved_tensors
        kernel, p, ncrit, wavek, precision = ctx.args
        dq, dy = _adj.adjoint(kernel, y.detach().cpu().double().numpy(), q.detach().cpu().double().numpy(),
                              x.detach().cpu().double().numpy(), gphi.detach().cpu().double().numpy(), p, ncrit,
                              wavek, precision)
        gq = torch.from_numpy(dq).to(q.dtype) if ctx.needs_input_grad[0] else None
        gy = torch.from_numpy(dy).to(y.dtype) if ctx.needs_input_grad[1] else None
        return gq, gy, None, None, None, None, None, None


def fmm_potential(q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
    """Differentiable phi at x for charges q at y (see FMMPotential)."""
    return FMMPotential.apply(q, y, x, kernel, p, ncrit, wavek, precision)
