"""torch.autograd.Function over the GPU evaluator: phi = FMM(q, y; x) with a hand-written backward that
runs the adjoint FMM pass (exafmm.adjoint) on the tree and operators built by the forward call.

Tensors are CPU tensors (float64, complex128 for Helmholtz): the evaluator runs on the GPU through its own
HIP runtime, and a process that also initialises torch's GPU runtime (any torch.cuda call) leaves the
device to that runtime, so the data crosses as host arrays.

Gradient convention (torch's, for complex tensors the conjugate Wirtinger one): with phi = A q,
A_ij = G(x_i, y_j), and g = dloss/dphi as torch passes it,
    grad_q = A^H g = conj(Phi_{conj g}(y)),    grad_y = Re(q * grad Phi_{conj g}(y)),
where Phi_c is the adjoint evaluation at the sources of charges c at the targets.  For the real kernels
the conjugations and Re are no-ops.
"""
import numpy as np
import torch

import exafmm
from exafmm.adjoint import AdjointFmm


class FMMPotential(torch.autograd.Function):
    """phi (ntrg,) at targets x (ntrg, 3) of charges q (nsrc,) at sources y (nsrc, 3).

    forward(q, y, x, kernel, p, ncrit, wavek, precision); gradients for q and y (x is treated as data)."""

    @staticmethod
    def forward(ctx, q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
        eng = AdjointFmm(kernel, y.detach().cpu().double().numpy(), x.detach().cpu().double().numpy(), p, ncrit,
                         wavek, precision)
        qn = q.detach().cpu().numpy().astype(eng.dtype)
        v = eng(qn)
        ctx.engine = eng
        ctx.q = qn
        ctx.dtypes = (q.dtype, y.dtype)
        return torch.from_numpy(v[:, 0].copy()).to(q.dtype)

    @staticmethod
    def backward(ctx, gphi):
        eng = ctx.engine
        g = gphi.detach().cpu().numpy().astype(eng.dtype)
        gq, gy = torch_grads(ctx.q, eng.adjoint_values(np.conj(g)))
        gq = torch.from_numpy(gq).to(ctx.dtypes[0]) if ctx.needs_input_grad[0] else None
        gy = torch.from_numpy(gy).to(ctx.dtypes[1]) if ctx.needs_input_grad[1] else None
        return gq, gy, None, None, None, None, None, None


def torch_grads(q, v):
    """(grad_q, grad_y) in torch's convention from the adjoint values v (nsrc, 4) = (Phi, grad Phi) at the
    sources of the charges conj(g) at the targets: conj(Phi) and Re(q grad Phi)."""
    return np.conj(v[:, 0]), np.ascontiguousarray(np.real(q[:, None] * v[:, 1:4]))


def fmm_potential(q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
    """Differentiable phi at x for charges q at y (see FMMPotential)."""
    return FMMPotential.apply(q, y, x, kernel, p, ncrit, wavek, precision)
