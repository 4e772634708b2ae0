"""Adjoint FMM pass: derivatives of L = sum_i w_i phi_i with respect to the charges and source positions.

With phi_i = sum_j q_j G(x_i, y_j) (targets x, sources y), the adjoint pass places w as charges at the
targets and evaluates potential and gradient at the sources:

    dL/dq_j = Phi_w(y_j),        dL/dy_j = q_j grad Phi_w(y_j),     Phi_w(t) = sum_i w_i G(t, x_i),

because G depends on |t - s| only, so G(x, y) = G(y, x) and grad_y G(x, y) = grad_t G(t, x) at t = y.
This holds for all three kernels of the reference (laplace.h:57, helmholtz.h, modified_helmholtz.h
gradient_P2P); for Helmholtz the charges, w and L are complex and both derivatives are the holomorphic
ones.  Pairs with r = 0 are excluded in both passes, as in the forward evaluation.

`AdjointFmm` builds one tree and one operator set and serves both passes from it: the tree is built over
the sources when the targets are the same points (the usual N-body case), otherwise over the union of
sources and targets, with zero charges on the points that do not act in a pass (forward: charges q on
the source part, zero on the target part; adjoint: zero on the source part, w on the target part).
Both passes run the GPU evaluator; nothing here computes on the CPU but the host setup.
"""
import os
import tempfile

import numpy as np

import exafmm

KERNELS = ("laplace", "helmholtz", "modified_helmholtz")
_CACHE = None


def _default_cache(kernel, p, wavek):
    """Operator cache file in a private temporary directory (the kernels' own default is a file in the
    working directory)."""
    global _CACHE
    if _CACHE is None:
        _CACHE = tempfile.mkdtemp(prefix="exafmm_adjoint_")
    return os.path.join(_CACHE, "%s_p%d_k%s.dat" % (kernel, p, str(wavek).replace(" ", "")))


def _fmm(kernel, p, ncrit, wavek, precision, filename):
    if kernel not in KERNELS:
        raise ValueError("adjoint: kernel must be one of %s, got %r" % (KERNELS, kernel))
    filename = filename or _default_cache(kernel, p, wavek)
    mod = getattr(exafmm, kernel)
    if kernel == "laplace":
        fmm = mod.LaplaceFmm(p, ncrit, filename)
    elif kernel == "helmholtz":
        fmm = mod.HelmholtzFmm(p, ncrit, complex(wavek), filename)
    else:
        fmm = mod.ModifiedHelmholtzFmm(p, ncrit, float(wavek), filename)
    fmm.precision = precision
    return mod, fmm


class AdjointFmm:
    """One tree and operator set for the forward evaluation phi(x) of charges q at y and for the adjoint
    evaluation at y of charges w at x.

    src (nsrc, 3), trg (ntrg, 3): positions, fixed for the life of the object (new positions need a new
    object: the tree depends on them).  `forward(q)` -> (ntrg, 4) potential and gradient at the targets;
    `backward(w)` -> (dL/dq (nsrc,), dL/dy (nsrc, 3)) for L = sum_i w_i phi_i.  Charges are float64 for
    the real kernels and complex128 for Helmholtz.  `filename` caches the operators in the reference's
    .dat layout (default: a file in a private temporary directory)."""

    def __init__(self, kernel, src, trg, p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64, filename=""):
        self.kernel = kernel
        self.dtype = np.complex128 if kernel == "helmholtz" else np.float64
        src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
        trg = np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
        if src.shape[0] == 0 or trg.shape[0] == 0:
            raise ValueError("AdjointFmm: sources and targets must be non-empty")
        self.nsrc, self.ntrg = src.shape[0], trg.shape[0]
        self.same = src.shape == trg.shape and np.array_equal(src, trg)
        pts = src if self.same else np.concatenate([src, trg])
        self.mod, self.fmm = _fmm(kernel, p, ncrit, wavek, precision, filename)
        self.tree = self.mod.setup(self.mod.init_sources(pts, np.zeros(pts.shape[0], self.dtype)),
                                   self.mod.init_targets(pts), self.fmm)

    def _charges(self, part, c):
        """Charges of the tree's source bodies: c on `part` ("src" or "trg"), zero elsewhere."""
        if self.same:
            return c
        full = np.zeros(self.nsrc + self.ntrg, self.dtype)
        if part == "src":
            full[:self.nsrc] = c
        else:
            full[self.nsrc:] = c
        return full

    def _run(self, part, c):
        n = self.nsrc if part == "src" else self.ntrg
        c = np.ascontiguousarray(c, dtype=self.dtype).reshape(-1)
        if c.shape[0] != n:
            raise ValueError("AdjointFmm: %d values for %d %s points" % (c.shape[0], n,
                                                                         "source" if part == "src" else "target"))
        self.charges = self._charges(part, c)
        self.mod.update_charges(self.tree, self.charges)
        self.mod.clear_values(self.tree)
        return np.asarray(self.mod.evaluate(self.tree, self.fmm))

    def forward(self, q):
        """Potential and gradient (ntrg, 4) at the targets of charges q at the sources."""
        v = self._run("src", q)
        return v if self.same else v[self.nsrc:]

    def backward(self, w):
        """(dL/dq, dL/dy) for L = sum_i w_i phi_i(q) at the charges q of the last `forward` call.  dL/dq
        does not depend on q; dL/dy = q_j grad Phi_w(y_j) uses the last forward's q."""
        q = self.last_q
        v = self._run("trg", w)
        v = v if self.same else v[:self.nsrc]
        return v[:, 0].copy(), q[:, None] * v[:, 1:4]

    def adjoint_values(self, w):
        """Phi_w and grad Phi_w at the sources (nsrc, 4): the adjoint evaluation itself."""
        v = self._run("trg", w)
        return v if self.same else v[:self.nsrc]

    def __call__(self, q):
        self.last_q = np.ascontiguousarray(q, dtype=self.dtype).reshape(-1)
        return self.forward(self.last_q)

    def flat_tree(self):
        """The flattened tree the GPU evaluator runs (exafmm Fmm.flat_tree)."""
        return self.fmm.flat_tree(self.tree)

    def operators(self):
        return self.fmm.operators()

    def sorted_charges(self):
        """Charges of the last pass in the flattened tree's source order."""
        return self.fmm.sorted_charges(self.tree)

    def close(self):
        self.fmm.release_device()


def evaluate(kernel, src, q, trg, p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64, filename=""):
    """Potential and gradient at trg (ntrg, 3) of charges q at src (nsrc, 3): array (ntrg, 4), columns phi,
    dphi/dx, dphi/dy, dphi/dz (gradient with respect to the target position)."""
    a = AdjointFmm(kernel, src, trg, p, ncrit, wavek, precision, filename)
    try:
        return a(q)
    finally:
        a.close()


def adjoint(kernel, src, q, trg, w, p=8, ncrit=64, wavek=0.0, precision=exafmm.FP64, filename=""):
    """(dL/dq (nsrc,), dL/dsrc (nsrc, 3)) for L = sum_i w_i phi_i, phi = evaluate(...)[:, 0]; one tree for
    both passes."""
    a = AdjointFmm(kernel, src, trg, p, ncrit, wavek, precision, filename)
    try:
        a.last_q = np.ascontiguousarray(q, dtype=a.dtype).reshape(-1)
        if a.last_q.shape[0] != a.nsrc:
            raise ValueError("adjoint: %d charges for %d sources" % (a.last_q.shape[0], a.nsrc))
        return a.backward(w)
    finally:
        a.close()
