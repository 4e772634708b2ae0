"""Adjoint FMM pass (exafmm.adjoint) and the torch.autograd.Function over it (exafmm.autograd).

CPU: the adjoint formula itself (dL/dq = Phi_w(y), dL/dy = q grad Phi_w(y)) against central finite
differences of L computed with the oracle's long-double direct sum.  GPU: the adjoint FMM pass against
the same formula evaluated with the engine's FP64 direct sum, against central differences (h = 1e-4)
on 32 random coordinates, and through torch autograd.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "exafmm-t_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from exafmm import abi  # noqa: E402

KID = {"laplace": 0, "modified_helmholtz": 2}


def _sphere(n, rng):
    x = rng.normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1)[:, None]


@pytest.mark.parametrize("kernel,wavek", [("laplace", 0.0), ("modified_helmholtz", 5.0)])
def test_adjoint_formula_vs_finite_differences(kernel, wavek):
    rng = np.random.default_rng(1)
    y = rng.random((40, 3))            # sources
    x = rng.random((30, 3)) + 0.05     # targets (distinct from the sources)
    q = rng.normal(size=40)
    w = rng.normal(size=30)
    kid = KID[kernel]

    def L(yy, qq):
        return float(np.dot(w, O.direct(kid, yy, qq, x, wavek, gradient=False)))

    v = O.direct(kid, x, w, y, wavek, gradient=True)  # charges w at the targets, evaluated at the sources
    dq, dy = v[:, 0], q[:, None] * v[:, 1:4]
    h = 1e-5
    for j in range(0, 40, 7):
        e = np.zeros(40)
        e[j] = h
        fd = (L(y, q + e) - L(y, q - e)) / (2 * h)
        assert abs(fd - dq[j]) <= 1e-8 * (1 + abs(dq[j])), (j, fd, dq[j])
        for d in range(3):
            yp, ym = y.copy(), y.copy()
            yp[j, d] += h
            ym[j, d] -= h
            fd = (L(yp, q) - L(ym, q)) / (2 * h)
            assert abs(fd - dy[j, d]) <= 1e-6 * (1 + abs(dy[j, d])), (j, d, fd, dy[j, d])


def _case(n=3000, seed=3):
    rng = np.random.default_rng(seed)
    y = _sphere(n, rng)
    x = _sphere(n, rng)
    return y, x, rng.normal(size=n), rng.normal(size=n)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,wavek", [("laplace", 0.0), ("modified_helmholtz", 5.0)])
def test_adjoint_fmm_vs_direct(require_gpu, kernel, wavek):
    from exafmm import adjoint as A
    y, x, q, w = _case()
    dq, dy = A.adjoint(kernel, y, q, x, w, p=8, ncrit=64, wavek=wavek)
    v = abi.direct(KID[kernel], x, w, y, wavek, abi.FP64, True)
    eq = np.linalg.norm(dq - v[:, 0]) / np.linalg.norm(v[:, 0])
    ey = np.linalg.norm(dy - q[:, None] * v[:, 1:4]) / np.linalg.norm(q[:, None] * v[:, 1:4])
    print(kernel, "adjoint vs direct: dq %.2e dy %.2e" % (eq, ey))
    assert eq < 1e-6 and ey < 1e-5


@pytest.mark.gpu
def test_adjoint_central_differences(require_gpu):
    """dL/dq and dL/dy on 32 random coordinates vs central differences, h = 1e-4, of L computed with the
    engine's FP64 direct sum (Yukawa k = 5, sphere)."""
    from exafmm import adjoint as A
    y, x, q, w = _case(2000, seed=4)
    wavek = 5.0
    dq, dy = A.adjoint("modified_helmholtz", y, q, x, w, p=8, ncrit=64, wavek=wavek)

    def L(yy, qq):
        return float(np.dot(w, abi.direct(2, yy, qq, x, wavek, abi.FP64, False)))

    rng = np.random.default_rng(9)
    h = 1e-4
    scale_q, scale_y = np.abs(dq).max(), np.abs(dy).max()
    for _ in range(32):
        j, d = int(rng.integers(0, y.shape[0])), int(rng.integers(0, 4))
        if d == 3:
            e = np.zeros_like(q)
            e[j] = h
            fd = (L(y, q + e) - L(y, q - e)) / (2 * h)
            assert abs(fd - dq[j]) <= 1e-6 * scale_q, (j, fd, dq[j])
        else:
            yp, ym = y.copy(), y.copy()
            yp[j, d] += h
            ym[j, d] -= h
            fd = (L(yp, q) - L(ym, q)) / (2 * h)
            assert abs(fd - dy[j, d]) <= 1e-5 * scale_y, (j, d, fd, dy[j, d])


@pytest.mark.gpu
def test_torch_autograd(require_gpu):
    torch = pytest.importorskip("torch")
    from exafmm import adjoint as A
    from exafmm.autograd import fmm_potential
    y, x, q, w = _case(2500, seed=6)
    qt = torch.tensor(q, requires_grad=True)
    yt = torch.tensor(y, requires_grad=True)
    phi = fmm_potential(qt, yt, torch.tensor(x), "modified_helmholtz", 8, 64, 5.0)
    (phi * torch.tensor(w)).sum().backward()
    ref = A.evaluate("modified_helmholtz", y, q, x, 8, 64, 5.0)
    assert np.array_equal(phi.detach().numpy(), ref[:, 0])
    dq, dy = A.adjoint("modified_helmholtz", y, q, x, w, 8, 64, 5.0)
    assert np.array_equal(qt.grad.numpy(), dq)
    assert np.array_equal(yt.grad.numpy(), dy)
