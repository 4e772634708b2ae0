"""ctypes wrapper of the oracle (oracle.c).  TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker of the gfx950 engine; the product never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
LAPLACE, HELMHOLTZ, YUKAWA = 0, 1, 2
ST_P2M, ST_M2M, ST_P2L, ST_M2P, ST_P2P, ST_M2L, ST_L2L, ST_L2P = 1, 2, 4, 8, 16, 32, 64, 128
ST_UP = ST_P2M | ST_M2M
ST_DOWN = ST_P2L | ST_M2P | ST_P2P | ST_M2L | ST_L2L | ST_L2P
STAGES = {"P2M": ST_P2M, "M2M": ST_M2M, "P2L": ST_P2L, "M2P": ST_M2P, "P2P": ST_P2P, "M2L": ST_M2L,
          "L2L": ST_L2L, "L2P": ST_L2P}

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class OrcTree(ctypes.Structure):
    _fields_ = [
        ("nnodes", ctypes.c_int), ("depth", ctypes.c_int), ("r0", ctypes.c_double), ("x0", ctypes.c_double * 3),
        ("level_offset", _c_int_p), ("center", _c_dbl_p), ("child0", _c_int_p),
        ("src_begin", _c_int_p), ("src_count", _c_int_p), ("trg_begin", _c_int_p), ("trg_count", _c_int_p),
        ("nsrcs", ctypes.c_int), ("ntrgs", ctypes.c_int), ("src_coord", _c_dbl_p), ("trg_coord", _c_dbl_p),
        ("p2p_offset", _c_int_p), ("p2p_source", _c_int_p), ("m2p_offset", _c_int_p), ("m2p_source", _c_int_p),
        ("p2l_offset", _c_int_p), ("p2l_source", _c_int_p), ("m2l_source", _c_int_p),
    ]


class OrcOps(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("p", ctypes.c_int), ("kre", ctypes.c_double), ("kim", ctypes.c_double),
        ("nlevels", ctypes.c_int), ("nm2l", ctypes.c_int),
        ("uc2e_u", _c_dbl_p), ("uc2e_v", _c_dbl_p), ("dc2e_u", _c_dbl_p), ("dc2e_v", _c_dbl_p),
        ("m2m", _c_dbl_p), ("l2l", _c_dbl_p), ("m2l", _c_dbl_p),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_direct.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp, ctypes.c_int, _vp,
                                   ctypes.c_int, _vp, ctypes.c_int]
        lib.orc_surface.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, ctypes.c_double, _vp]
        lib.orc_surf2conv.argtypes = [ctypes.c_int, ctypes.c_int, _vp]
        lib.orc_dft3_r2c.argtypes = [_vp, ctypes.c_int, _vp]
        lib.orc_dft3_c2r.argtypes = [_vp, ctypes.c_int, _vp]
        lib.orc_dft3_c2c.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
        lib.orc_m2l_operator.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp]
        lib.orc_evaluate.argtypes = [ctypes.POINTER(OrcTree), ctypes.POINTER(OrcOps), _vp, ctypes.c_int, _vp, _vp,
                                     _vp]
        lib.orc_fft.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp]
        lib.orc_evaluate_fast.argtypes = [ctypes.POINTER(OrcTree), ctypes.POINTER(OrcOps), _vp, _vp, _vp, _vp, _vp]
        lib.orc_evaluate_fast.restype = ctypes.c_int
        for f in ("orc_direct", "orc_surface", "orc_surf2conv", "orc_dft3_r2c", "orc_dft3_c2r", "orc_dft3_c2c",
                  "orc_m2l_operator", "orc_evaluate", "orc_fft"):
            getattr(lib, f).restype = None
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp)


def _kind(kernel):
    """engine kernel ids (0 Laplace, 1 Helmholtz, 2 modified Helmholtz) == oracle ids"""
    return int(kernel)


def direct(kernel, src, q, trg, wavek=0.0, gradient=True):
    lib = load()
    cplx = kernel == HELMHOLTZ
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
    trg = np.ascontiguousarray(trg, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, dtype=np.complex128 if cplx else np.float64)
    out = np.zeros((trg.shape[0], 4) if gradient else (trg.shape[0],), dtype=q.dtype)
    wk = complex(wavek)
    lib.orc_direct(_kind(kernel), wk.real, wk.imag, _p(src), _p(q.view(np.float64)), src.shape[0], _p(trg),
                   trg.shape[0], _p(out.view(np.float64)), 1 if gradient else 0)
    return out


def fft(x, sign):
    """Mixed-radix FFT of fast.c: out[k] = sum_j x[j] exp(sign 2 pi i j k / n), n <= 64."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib.orc_fft(x.shape[0], int(sign), _p(x.view(np.float64)), _p(out.view(np.float64)))
    return out


def surface(p, r0, level, c, alpha):
    lib = load()
    n = 6 * (p - 1) ** 2 + 2
    out = np.zeros((n, 3))
    c = np.ascontiguousarray(c, dtype=np.float64)
    lib.orc_surface(p, r0, level, _p(c), alpha, _p(out))
    return out


def surf2conv(p, dn):
    lib = load()
    m = np.zeros(6 * (p - 1) ** 2 + 2, dtype=np.int32)
    lib.orc_surf2conv(p, 1 if dn else 0, _p(m))
    return m


def dft3_r2c(x):
    lib = load()
    n = x.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    lib.orc_dft3_r2c(_p(x), n, _p(out.view(np.float64)))
    return out


def dft3_c2r(X, n):
    lib = load()
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros((n, n, n))
    lib.orc_dft3_c2r(_p(X.view(np.float64)), n, _p(out))
    return out


def dft3_c2c(X, sign):
    lib = load()
    n = X.shape[0]
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros_like(X)
    lib.orc_dft3_c2c(_p(X.view(np.float64)), n, sign, _p(out.view(np.float64)))
    return out


def m2l_operator(kernel, p, r0, child_level, scale_invariant, wavek=0.0):
    lib = load()
    n1 = 2 * p
    nfreq = n1 ** 3 if kernel == HELMHOLTZ else n1 * n1 * (n1 // 2 + 1)
    out = np.zeros(26 * nfreq * 128)
    wk = complex(wavek)
    lib.orc_m2l_operator(_kind(kernel), wk.real, wk.imag, p, r0, child_level, 1 if scale_invariant else 0, _p(out))
    return out


class Oracle:
    """Oracle evaluation over a flattened tree + operators (same arrays as the engine)."""

    def __init__(self, kernel, p, flat, ops, wavek=0.0):
        self.lib = load()
        self.kernel, self.p = kernel, p
        self.cplx = kernel == HELMHOLTZ
        self.nsurf = 6 * (p - 1) ** 2 + 2
        self.keep = {}
        t = OrcTree()
        t.nnodes, t.depth, t.r0 = int(flat["nnodes"]), int(flat["depth"]), float(flat["r0"])
        for k in ("level_offset", "child0", "src_begin", "src_count", "trg_begin", "trg_count", "p2p_offset",
                  "p2p_source", "m2p_offset", "m2p_source", "p2l_offset", "p2l_source", "m2l_source"):
            a = np.ascontiguousarray(flat[k], dtype=np.int32)
            self.keep[k] = a
            setattr(t, k, a.ctypes.data_as(_c_int_p))
        for k in ("center", "src_coord", "trg_coord"):
            a = np.ascontiguousarray(flat[k], dtype=np.float64)
            self.keep[k] = a
            setattr(t, k, a.ctypes.data_as(_c_dbl_p))
        t.nsrcs = self.keep["src_coord"].size // 3
        t.ntrgs = self.keep["trg_coord"].size // 3
        o = OrcOps()
        wk = complex(wavek)
        o.kind, o.p, o.kre, o.kim = _kind(kernel), p, wk.real, wk.imag
        o.nlevels, o.nm2l = int(ops["nlevels"]), int(ops["nm2l"])
        for k in ("uc2e_u", "uc2e_v", "dc2e_u", "dc2e_v", "m2m", "l2l", "m2l"):
            a = np.ascontiguousarray(ops[k])
            a = a.view(np.float64) if np.iscomplexobj(a) else a.astype(np.float64, copy=False)
            self.keep["op_" + k] = a
            setattr(o, k, a.ctypes.data_as(_c_dbl_p))
        self.t, self.o = t, o
        self.nnodes, self.ntrgs = t.nnodes, t.ntrgs
        dt = np.complex128 if self.cplx else np.float64
        self.up = np.zeros((t.nnodes, self.nsurf), dtype=dt)
        self.dn = np.zeros((t.nnodes, self.nsurf), dtype=dt)
        self.val = np.zeros((t.ntrgs, 4), dtype=dt)
        self.q = np.zeros(t.nsrcs, dtype=dt)

    def set_charges(self, q):
        self.q = np.ascontiguousarray(q, dtype=np.complex128 if self.cplx else np.float64)

    def clear(self):
        self.up[:] = 0
        self.dn[:] = 0
        self.val[:] = 0

    def run(self, stages):
        self.lib.orc_evaluate(ctypes.byref(self.t), ctypes.byref(self.o), _p(self.q.view(np.float64)), int(stages),
                              _p(self.up.view(np.float64)), _p(self.dn.view(np.float64)),
                              _p(self.val.view(np.float64)))

    def evaluate(self):
        self.clear()
        self.run(ST_UP)
        self.run(ST_DOWN)
        return self.val

    FAST_STAGES = ("setup", "P2M", "M2M", "P2L+M2P", "P2P", "FFT", "HAD", "IFFT", "L2L", "L2P")

    def evaluate_fast(self):
        """The FFT-based CPU evaluation of fast.c (Laplace, one operator level): same result as evaluate().
        Seconds per stage in self.fast_stage_s."""
        if self.kernel != LAPLACE:
            raise ValueError("evaluate_fast: Laplace only")
        st = np.zeros(len(self.FAST_STAGES))
        rc = self.lib.orc_evaluate_fast(ctypes.byref(self.t), ctypes.byref(self.o), _p(self.q), _p(self.up),
                                        _p(self.dn), _p(self.val), _p(st))
        if rc != 0:
            raise ValueError("evaluate_fast: unsupported operators")
        self.fast_stage_s = dict(zip(self.FAST_STAGES, st.tolist()))
        return self.val
