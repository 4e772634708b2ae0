"""torch.autograd.Function over the GPU evaluator: phi = FMM(q, y; x) with a hand-written backward that
runs the adjoint FMM pass (exafmm.adjoint) on the tree and operators built by the forward call.

Tensors (float64, complex128 for Helmholtz) may live on the CPU or on the GPU; outputs and gradients come
back on the device of the charges / source coordinates.  The octree is built on the host from the
coordinates, so positions and charges cross to host memory once per call (N x 32 bytes) and the results
cross back.  Device tensors need one HIP runtime in the process: torch's wheel ships its own
libamdhip64 with the same soname as the system one, and the engine binds to whichever is loaded first, so
`import torch` before the first `import exafmm` (then the engine runs on torch's runtime and both see the
same device).  With two runtimes loaded, GPU tensors are refused with an explanation (CPU tensors work
either way).

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


def hip_runtimes():
    """Paths of the HIP runtime libraries (libamdhip64) mapped into this process."""
    out = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 6 and "libamdhip64" in parts[-1]:
                    out.add(parts[-1])
    except OSError:
        pass
    return sorted(out)


def _check_device(*tensors):
    if any(t.is_cuda for t in tensors):
        libs = hip_runtimes()
        if len(libs) > 1:
            raise RuntimeError(
                "exafmm.autograd: GPU tensors need the engine and torch on one HIP runtime, but this process "
                "has %d loaded (%s): import torch before exafmm" % (len(libs), ", ".join(libs)))


class FMMPotential(torch.autograd.Function):
    """phi (ntrg,) at targets x (ntrg, 3) of charges q (nsrc,) at sources y (nsrc, 3).

    forward(q, y, x, kernel, p, ncrit, wavek, precision); gradients for q and y (x is treated as data)."""

    @staticmethod
    def forward(ctx, q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
        _check_device(q, y, x)
        eng = AdjointFmm(kernel, y.detach().cpu().double().numpy(), x.detach().cpu().double().numpy(), p, ncrit,
                         wavek, precision)
        qn = q.detach().cpu().numpy().astype(eng.dtype)
        v = eng(qn)
        ctx.engine = eng
        ctx.q = qn
        ctx.dtypes = (q.dtype, y.dtype)
        ctx.devices = (q.device, y.device)
        return torch.from_numpy(v[:, 0].copy()).to(device=q.device, dtype=q.dtype)

    @staticmethod
    def backward(ctx, gphi):
        eng = ctx.engine
        g = gphi.detach().cpu().numpy().astype(eng.dtype)
        gq, gy = torch_grads(ctx.q, eng.adjoint_values(np.conj(g)))
        gq = torch.from_numpy(gq).to(device=ctx.devices[0], dtype=ctx.dtypes[0]) if ctx.needs_input_grad[0] else None
        gy = torch.from_numpy(gy).to(device=ctx.devices[1], dtype=ctx.dtypes[1]) if ctx.needs_input_grad[1] else None
        return gq, gy, None, None, None, None, None, None


def torch_grads(q, v):
    """(grad_q, grad_y) in torch's convention from the adjoint values v (nsrc, 4) = (Phi, grad Phi) at the
    sources of the charges conj(g) at the targets: conj(Phi) and Re(q grad Phi)."""
    return np.conj(v[:, 0]), np.ascontiguousarray(np.real(q[:, None] * v[:, 1:4]))


def fmm_potential(q, y, x, kernel="laplace", p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64):
    """Differentiable phi at x for charges q at y (see FMMPotential)."""
    return FMMPotential.apply(q, y, x, kernel, p, ncrit, wavek, precision)
