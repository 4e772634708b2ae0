"""xfmm_direct_noalloc (the allocation-free direct sum) against xfmm_direct.

GPU: inputs and outputs live in device memory allocated here by the test
through the HIP runtime the engine links; the entry point itself allocates
nothing.  Two calls in a row on one stream, the second with a smaller input,
each compared with xfmm_direct on the same data (FP64: 1e-13 relative, the
summation order is the same; mixed: 1e-6, the FP32 direct-sum bar of the
reference's tests/p2p_laplace.cpp:82-84).  The output buffer starts non-zero
to check the += of the reference's gradient_P2P.
CPU: sizes beyond the scratch capacity are rejected before anything runs.
"""
import ctypes

import numpy as np
import pytest

from exafmm import abi

H2D, D2H = 1, 2
SCRATCH_BYTES = 32 << 20  # XFMM_NOALLOC_SCRATCH_BYTES


class Hip:
    """Device memory helpers from the HIP runtime that the engine library links."""

    def __init__(self):
        abi.load()
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [vp]
        self.lib.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.lib.hipStreamDestroy.argtypes = [vp]
        self.lib.hipStreamSynchronize.argtypes = [vp]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, "%s failed with HIP error %d" % (what, rc)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.bufs.append(p)
        self.check(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        self.check(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def noalloc_lib():
    lib = abi.load()
    f = lib.xfmm_direct_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return lib


def case(rng, kernel, ns, nt):
    src = rng.random((ns, 3))
    trg = np.concatenate([src[: nt // 2], rng.random((nt - nt // 2, 3))])  # coincident pairs included
    if kernel == abi.HELMHOLTZ:
        q = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
        init = (rng.random((nt, 4)) + 1j * rng.random((nt, 4)))
    else:
        q = rng.uniform(-1, 1, ns)
        init = rng.random((nt, 4))
    return src, q, trg, init


def test_direct_noalloc_rejects_sizes_beyond_scratch():
    lib = noalloc_lib()
    buf = np.zeros(16)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    for kernel, na in ((abi.LAPLACE, 4), (abi.HELMHOLTZ, 8)):
        nt = SCRATCH_BYTES // 8 // na + 1
        rc = lib.xfmm_direct_noalloc(kernel, abi.FP64, 1.0, 0.0, p, p, 1, p, nt, p, 1, None)
        assert rc == 1
        assert b"scratch" in lib.xfmm_last_error(None)
    rc = lib.xfmm_direct_noalloc(abi.LAPLACE, abi.FP64, 0.0, 0.0, p, p, -1, p, 1, p, 1, None)
    assert rc == 1


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,wavek,precision", [
    (abi.LAPLACE, 0.0, abi.FP64), (abi.LAPLACE, 0.0, abi.MIXED),
    (abi.MODIFIED_HELMHOLTZ, 5.0, abi.FP64), (abi.HELMHOLTZ, 20.0 + 0.0j, abi.FP64)])
def test_direct_noalloc_matches_direct(require_gpu, kernel, wavek, precision):
    lib = noalloc_lib()
    hip = Hip()
    stream = ctypes.c_void_p()
    hip.check(hip.lib.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    rng = np.random.default_rng(11 + kernel)
    tol = 1e-13 if precision == abi.FP64 else 1e-6
    wk = complex(wavek)
    try:
        for ns, nt in ((6000, 4000), (700, 300)):  # the second call is smaller
            src, q, trg, init = case(rng, kernel, ns, nt)
            print("kernel %d precision %d ns %d nt %d" % (kernel, precision, ns, nt), flush=True)
            ref = abi.direct(kernel, src, q, trg, wavek, precision, True, out=init.copy())
            qv = q.view(np.float64) if np.iscomplexobj(q) else q
            iv = init.view(np.float64)
            d_src, d_q, d_trg, d_out = hip.upload(src), hip.upload(qv), hip.upload(trg), hip.upload(iv)
            rc = lib.xfmm_direct_noalloc(kernel, precision, wk.real, wk.imag, d_src, d_q, ns, d_trg, nt, d_out, 1,
                                         stream)
            assert rc == 0, lib.xfmm_last_error(None).decode()
            hip.check(hip.lib.hipStreamSynchronize(stream), "hipStreamSynchronize")
            got = hip.download(d_out, iv.shape)
            if np.iscomplexobj(ref):
                got = got.view(np.complex128)
            err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref - init))
            print("  rel diff vs xfmm_direct %.3e" % err, flush=True)
            assert err <= tol, err
            hip.free()
    finally:
        hip.free()
        hip.lib.hipStreamDestroy(stream)
