"""Adjoint FMM pass (exafmm.adjoint) and the torch.autograd.Function over it (exafmm.autograd).

CPU:
* the adjoint formula (dL/dq = Phi_w(y), dL/dy = q grad Phi_w(y)) for all three kernels against central
  finite differences of L computed with the oracle's long-double direct sum (complex L for Helmholtz);
* the torch gradient convention of the backward (grad_q = conj(Phi_{conj g}), grad_y = Re(q grad
  Phi_{conj g})) against torch's own autograd through a direct sum written in torch (complex Helmholtz
  charges, a real loss).
GPU:
* the adjoint FMM pass against the same formula evaluated with the engine's FP64 direct sum (three
  kernels, one tree for both passes, sources = targets and distinct point sets);
* at the Yukawa workload's parameters (k = 5, unit sphere, p = 8, ncrit = 64, adaptive tree with
  non-empty M2P / P2L lists): the adjoint evaluation against the oracle on the same tree and operators,
  FP64 within 1e-10, mixed within 1e-5;
* central differences (h = 1e-4) on 32 random coordinates of L computed by the FMM itself with the tree
  held fixed (the flattened tree with one source coordinate moved), and of L computed by the direct sum;
* through torch.autograd (real and complex charges).
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
from fmm_cases import rel_err  # noqa: E402

KID = {"laplace": 0, "helmholtz": 1, "modified_helmholtz": 2}
WAVEK = {"laplace": 0.0, "helmholtz": 10.0 + 0.0j, "modified_helmholtz": 5.0}


def _sphere(n, rng):
    x = rng.normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1)[:, None]


def _charges(kernel, n, rng):
    if kernel == "helmholtz":
        return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    return rng.normal(size=n)


@pytest.mark.parametrize("kernel", ["laplace", "helmholtz", "modified_helmholtz"])
def test_adjoint_formula_vs_finite_differences(kernel):
    rng = np.random.default_rng(1)
    y = rng.random((40, 3))            # sources
    x = rng.random((30, 3)) + 0.05     # targets (distinct from the sources)
    q = _charges(kernel, 40, rng)
    w = _charges(kernel, 30, rng)
    kid, wavek = KID[kernel], WAVEK[kernel]

    def L(yy, qq):
        return np.dot(w, O.direct(kid, yy, qq, x, wavek, gradient=False))

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


def test_torch_convention_complex():
    """torch_grads applied to the adjoint of conj(g) equals torch's autograd through a torch direct sum."""
    torch = pytest.importorskip("torch")
    from exafmm.autograd import torch_grads
    rng = np.random.default_rng(2)
    y = rng.random((25, 3))
    x = rng.random((20, 3)) + 0.05
    q = _charges("helmholtz", 25, rng)
    k = 10.0

    qt = torch.tensor(q, requires_grad=True)
    yt = torch.tensor(y, requires_grad=True)
    d = torch.tensor(x)[:, None, :] - yt[None, :, :]
    r = torch.sqrt((d * d).sum(-1))
    G = torch.exp(1j * k * r) / (4 * np.pi * r)
    phi = G @ qt
    c = torch.tensor(rng.normal(size=20) + 1j * rng.normal(size=20))
    loss = (phi * c).real.sum() + (phi.abs() ** 2).sum()
    loss.backward()
    g = (c.conj() + 2 * phi).detach().numpy()  # dloss/dphi in torch's convention
    v = O.direct(1, x, np.conj(g), y, k + 0j, gradient=True)
    gq, gy = torch_grads(q, v)
    assert np.allclose(gq, qt.grad.numpy(), rtol=1e-10, atol=1e-12)
    assert np.allclose(gy, yt.grad.numpy(), rtol=1e-10, atol=1e-12)


def _case(kernel, n=3000, seed=3, distinct=False):
    rng = np.random.default_rng(seed)
    y = _sphere(n, rng)
    x = _sphere(n, rng) if distinct else y
    return y, x, _charges(kernel, n, rng), _charges(kernel, n, rng)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["laplace", "helmholtz", "modified_helmholtz"])
@pytest.mark.parametrize("distinct", [False, True])
def test_adjoint_fmm_vs_direct(require_gpu, kernel, distinct):
    from exafmm.adjoint import AdjointFmm
    y, x, q, w = _case(kernel, distinct=distinct)
    wavek = WAVEK[kernel]
    a = AdjointFmm(kernel, y, x, p=8, ncrit=64, wavek=wavek)
    phi = a(q)
    dq, dy = a.backward(w)
    phi2 = a(q)  # the tree serves the forward again after the adjoint pass
    a.close()
    assert np.array_equal(phi, phi2)
    ref = abi.direct(KID[kernel], y, q, x, wavek, abi.FP64, True)
    v = abi.direct(KID[kernel], x, w, y, wavek, abi.FP64, True)
    ef = rel_err(phi[:, 0], ref[:, 0])
    eq = rel_err(dq, v[:, 0])
    ey = rel_err(dy, q[:, None] * v[:, 1:4])
    print("%s distinct=%s: forward %.2e adjoint dq %.2e dy %.2e" % (kernel, distinct, ef, eq, ey), flush=True)
    assert ef < 1e-6 and eq < 1e-6 and ey < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("distinct", [False, True])
def test_adjoint_vs_oracle_yukawa_sphere(require_gpu, distinct):
    """The adjoint evaluation (w at the targets, values at the sources) at the Yukawa workload's
    parameters against the oracle on the same tree and operators: FP64 1e-10, mixed 1e-5 (measured
    7.9e-8 / 9.1e-10 with sources = targets, 9.1e-8 / 6.3e-9 for distinct sets)."""
    from exafmm.adjoint import AdjointFmm
    n = 40000 if not distinct else 20000
    y, x, q, w = _case("modified_helmholtz", n, seed=11, distinct=distinct)
    for prec, tol, gtol in ((abi.FP64, 1e-10, 1e-10), (abi.MIXED, 1e-5, 1e-5)):
        a = AdjointFmm("modified_helmholtz", y, x, p=8, ncrit=64, wavek=5.0, precision=prec)
        flat = a.flat_tree()
        m2p, p2l = int(flat["m2p_offset"][-1]), int(flat["p2l_offset"][-1])
        a.last_q = q
        got = a.adjoint_values(w)
        orc = O.Oracle(O.YUKAWA, 8, flat, a.operators(), 5.0)
        orc.set_charges(a.sorted_charges())
        ref_sorted = orc.evaluate()
        a.close()
        ref = np.empty_like(ref_sorted)
        ref[np.asarray(flat["trg_ibody"])] = ref_sorted
        ref = ref[:y.shape[0]]
        e = (rel_err(got[:, 0], ref[:, 0]), rel_err(got[:, 1:], ref[:, 1:]))
        print("yukawa sphere N=%d distinct=%s precision %d (m2p %d, p2l %d): adjoint vs oracle phi %.2e grad %.2e"
              % (n, distinct, prec, m2p, p2l, e[0], e[1]), flush=True)
        assert m2p > 0 and p2l > 0
        assert e[0] < tol and e[1] < gtol, (prec, e)


@pytest.mark.gpu
def test_adjoint_central_differences_fixed_tree(require_gpu):
    """dL/dq and dL/dy on 32 random coordinates against central differences (h = 1e-4) of L computed by
    the FMM with the tree held fixed (one source coordinate of the flattened tree moved), and of L
    computed by the engine's FP64 direct sum (Yukawa k = 5, distinct point sets on the sphere)."""
    from exafmm.adjoint import AdjointFmm
    y, x, q, w = _case("modified_helmholtz", 2000, seed=4, distinct=True)
    wavek = 5.0
    a = AdjointFmm("modified_helmholtz", y, x, p=8, ncrit=64, wavek=wavek)
    a(q)
    qs = a.sorted_charges()
    dq, dy = a.backward(w)
    flat, ops = a.flat_tree(), a.operators()
    a.close()
    sib, tib = np.asarray(flat["src_ibody"]), np.asarray(flat["trg_ibody"])
    wt = np.zeros(tib.shape[0])
    keep = tib >= y.shape[0]
    wt[keep] = w[tib[keep] - y.shape[0]]  # weights in the tree's target order (zero on the source part)
    spos = np.empty(sib.shape[0], dtype=np.int64)
    spos[sib] = np.arange(sib.shape[0])
    eng = abi.Engine(KID["modified_helmholtz"], 8, wavek, abi.FP64)
    eng.set_operators(ops)

    def L_fmm(coord, qq):
        f = dict(flat)
        f["src_coord"] = coord
        eng.set_tree(f)
        eng.set_charges(qq)
        eng.upward_pass()
        eng.downward_pass()
        return float(np.dot(wt, eng.values()[:, 0]))

    def L_direct(yy, qq):
        return float(np.dot(w, abi.direct(2, yy, qq, x, wavek, abi.FP64, False)))

    c0 = np.asarray(flat["src_coord"], dtype=np.float64).copy()
    rng = np.random.default_rng(9)
    h = 1e-4
    scale_q, scale_y = np.abs(dq).max(), np.abs(dy).max()
    worst = [0.0, 0.0]
    for _ in range(32):
        j, d = int(rng.integers(0, y.shape[0])), int(rng.integers(0, 4))
        s = int(spos[j])
        if d == 3:
            qp, qm = qs.copy(), qs.copy()
            qp[s] += h
            qm[s] -= h
            fd = (L_fmm(c0, qp) - L_fmm(c0, qm)) / (2 * h)
            e = np.zeros_like(q)
            e[j] = h
            fdd = (L_direct(y, q + e) - L_direct(y, q - e)) / (2 * h)
            got, scale = dq[j], scale_q
        else:
            cp, cm = c0.copy(), c0.copy()
            cp[3 * s + d] += h
            cm[3 * s + d] -= h
            fd = (L_fmm(cp, qs) - L_fmm(cm, qs)) / (2 * h)
            yp, ym = y.copy(), y.copy()
            yp[j, d] += h
            ym[j, d] -= h
            fdd = (L_direct(yp, q) - L_direct(ym, q)) / (2 * h)
            got, scale = dy[j, d], scale_y
        worst[0] = max(worst[0], abs(fd - got) / scale)
        worst[1] = max(worst[1], abs(fdd - got) / scale)
        assert abs(fd - got) <= 1e-5 * scale, (j, d, fd, got)
        assert abs(fdd - got) <= 1e-5 * scale, (j, d, fdd, got)
    eng.close()
    print("fixed-tree FMM central differences: max |fd - adjoint| / max|adjoint| = %.2e (direct-sum L: %.2e)"
          % tuple(worst), flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["modified_helmholtz", "helmholtz"])
def test_torch_autograd(require_gpu, kernel):
    torch = pytest.importorskip("torch")
    from exafmm.adjoint import AdjointFmm
    from exafmm.autograd import fmm_potential, torch_grads
    y, x, q, w = _case(kernel, 2500, seed=6)
    wavek = WAVEK[kernel]
    qt = torch.tensor(q, requires_grad=True)
    yt = torch.tensor(y, requires_grad=True)
    phi = fmm_potential(qt, yt, torch.tensor(x), kernel, 8, 64, wavek)
    wt = torch.tensor(w)
    loss = (phi * wt).real.sum()
    loss.backward()
    a = AdjointFmm(kernel, y, x, 8, 64, wavek)
    ref = a(q)
    assert np.array_equal(phi.detach().numpy(), ref[:, 0])
    # d loss / d phi in torch's convention is g = conj(w) for loss = Re(sum w phi); the backward runs the
    # adjoint evaluation of conj(g) = w
    gq, gy = torch_grads(q, a.adjoint_values(w))
    a.close()
    assert np.array_equal(qt.grad.numpy(), gq)
    assert np.array_equal(yt.grad.numpy(), gy)


_DEVICE_SCRIPT = r"""
import sys
import torch  # first: the engine then binds to the HIP runtime that torch loaded
sys.path.insert(0, sys.argv[1])
import numpy as np
from exafmm.autograd import fmm_potential, hip_runtimes

kernel, wavek = sys.argv[2], complex(sys.argv[3])
rng = np.random.default_rng(8)
n = 3000
y = rng.normal(size=(n, 3))
y /= np.linalg.norm(y, axis=1)[:, None]
x = rng.random((n, 3)) * 0.5
if kernel == "helmholtz":
    q = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    w = rng.normal(size=n) + 1j * rng.normal(size=n)
else:
    wavek = wavek.real
    q = rng.normal(size=n)
    w = rng.normal(size=n)
dev = torch.device("cuda")


def run(device):
    qt = torch.tensor(q, device=device, requires_grad=True)
    yt = torch.tensor(y, device=device, requires_grad=True)
    phi = fmm_potential(qt, yt, torch.tensor(x, device=device), kernel, 8, 64, wavek)
    loss = (phi * torch.tensor(w, device=device)).real.sum()
    loss.backward()
    return phi, qt.grad, yt.grad


phi_d, gq_d, gy_d = run(dev)
libs = hip_runtimes()
print("HIP runtimes in the process:", libs, flush=True)
assert len(libs) == 1, libs
assert phi_d.is_cuda and gq_d.is_cuda and gy_d.is_cuda
phi_c, gq_c, gy_c = run("cpu")
# same engine, same tree: the device path returns the CPU path's values
assert torch.equal(phi_d.cpu(), phi_c) and torch.equal(gq_d.cpu(), gq_c) and torch.equal(gy_d.cpu(), gy_c)
# torch's own GPU arithmetic beside the engine: a direct sum of the potential and of d loss / d q on the
# GPU in float64 (torch's autograd through the direct sum)
qt = torch.tensor(q, device=dev, requires_grad=True)
r = torch.cdist(torch.tensor(x, device=dev), torch.tensor(y, device=dev))
if kernel == "helmholtz":
    G = torch.exp(1j * wavek * r) / (4 * np.pi * r)
else:
    G = torch.exp(-wavek * r) / (4 * np.pi * r)
phi_t = G @ qt
(phi_t * torch.tensor(w, device=dev)).real.sum().backward()
e_phi = (torch.linalg.norm(phi_d - phi_t.detach()) / torch.linalg.norm(phi_t.detach())).item()
e_gq = (torch.linalg.norm(gq_d - qt.grad) / torch.linalg.norm(qt.grad)).item()
print("device tensors: phi vs torch direct %.2e, grad q %.2e" % (e_phi, e_gq), flush=True)
assert e_phi < 1e-6 and e_gq < 1e-6, (e_phi, e_gq)
print("DEVICE_OK", flush=True)
"""


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["modified_helmholtz", "helmholtz"])
def test_torch_autograd_device_tensors(require_gpu, kernel):
    """GPU tensors through the autograd function, in a fresh process that imports torch first (one HIP
    runtime): outputs and gradients come back on the GPU, equal to the CPU-tensor path bit for bit, and
    match torch's own float64 direct sum (and its autograd) computed on the GPU beside the engine."""
    pytest.importorskip("torch")
    import subprocess
    r = subprocess.run([sys.executable, "-c", _DEVICE_SCRIPT, os.path.join(ROOT, "exafmm-t_amd"), kernel,
                        str(WAVEK[kernel])], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0 and "DEVICE_OK" in r.stdout


def test_autograd_refuses_device_tensors_with_two_runtimes(monkeypatch):
    """With two HIP runtimes in the process, GPU tensors are refused with an explanation (checked on the
    CPU with a stand-in tensor and a patched runtime list)."""
    torch = pytest.importorskip("torch")
    from exafmm import autograd as ag

    class FakeCuda:
        is_cuda = True

    monkeypatch.setattr(ag, "hip_runtimes", lambda: ["/a/libamdhip64.so", "/b/libamdhip64.so.7"])
    with pytest.raises(RuntimeError, match="import torch before exafmm"):
        ag._check_device(torch.zeros(1), FakeCuda())
    monkeypatch.setattr(ag, "hip_runtimes", lambda: ["/a/libamdhip64.so"])
    ag._check_device(torch.zeros(1), FakeCuda())
