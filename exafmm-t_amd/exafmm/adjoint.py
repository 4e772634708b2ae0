"""Adjoint FMM pass for the real kernels (Laplace, modified Helmholtz).

For L = sum_i w_i phi_i with phi_i = sum_j q_j G(x_i, y_j) (targets x, sources y), the adjoint pass
places w as charges at the targets and evaluates potential and gradient at the sources:

    dL/dq_j = Phi_w(y_j),        dL/dy_j = q_j grad Phi_w(y_j),     Phi_w(t) = sum_i w_i G(t, x_i),

using that G depends on |t - s| only, so grad_s G(x, s) = grad_t G(t, x) at t = s.  Pairs with
r = 0 are excluded in both passes, as in the forward evaluation (laplace.h:57).  Both passes run the
GPU evaluator (exafmm.<kernel>.setup + evaluate); nothing here runs on the CPU but the host setup.
The Helmholtz kernel (complex charges, complex L) is not covered.
"""
import numpy as np

import exafmm

_KERNELS = {
    "laplace": ("laplace", "LaplaceFmm"),
    "modified_helmholtz": ("modified_helmholtz", "ModifiedHelmholtzFmm"),
}


def _fmm(kernel, p, ncrit, wavek, precision, filename):
    if kernel not in _KERNELS:
        raise ValueError("adjoint: kernel must be 'laplace' or 'modified_helmholtz', got %r" % (kernel,))
    modname, cls = _KERNELS[kernel]
    mod = getattr(exafmm, modname)
    if kernel == "laplace":
        fmm = getattr(mod, cls)(p, ncrit, filename)
    else:
        fmm = getattr(mod, cls)(p, ncrit, float(wavek), filename)
    fmm.precision = precision
    return mod, fmm


def evaluate(kernel, src, q, trg, p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64, filename=""):
    """Potential and gradient at trg (ntrg, 3) of real charges q at src (nsrc, 3): array (ntrg, 4),
    columns phi, dphi/dx, dphi/dy, dphi/dz (gradient with respect to the target position).  A fresh tree
    is built for every call.  `filename` caches the operators (reference .dat layout); keep it empty for
    the modified Helmholtz kernel when the points move, since its operators depend on the box size and
    the cache does not record it."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
    trg = np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    if q.shape[0] != src.shape[0]:
        raise ValueError("evaluate: %d charges for %d sources" % (q.shape[0], src.shape[0]))
    mod, fmm = _fmm(kernel, p, ncrit, wavek, precision, filename)
    tree = mod.setup(mod.init_sources(src, q), mod.init_targets(trg), fmm)
    return np.asarray(mod.evaluate(tree, fmm))


def adjoint(kernel, src, q, trg, w, p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64, filename=""):
    """(dL/dq (nsrc,), dL/dsrc (nsrc, 3)) for L = sum_i w_i phi_i, phi = evaluate(...)[:, 0]."""
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
    trg = np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
    if w.shape[0] != trg.shape[0]:
        raise ValueError("adjoint: %d weights for %d targets" % (w.shape[0], trg.shape[0]))
    v = evaluate(kernel, trg, w, src, p, ncrit, wavek, precision, filename)
    return v[:, 0].copy(), q[:, None] * v[:, 1:4]
