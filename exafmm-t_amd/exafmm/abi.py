"""ctypes binding of the C ABI in include/exafmm_t_amd.h.

This is the binding a caller without the C++ layer uses: it drives the gfx950
engine directly from the flattened tree and operators (for example the arrays
returned by ``Fmm.flat_tree()`` / ``Fmm.operators()`` of the pybind module).
The engine library must exist; nothing here falls back to CPU code.
"""
import ctypes
import os

import numpy as np

LAPLACE, HELMHOLTZ, MODIFIED_HELMHOLTZ = 0, 1, 2
FP64, MIXED, FLOAT = 0, 1, 2
TUNE_HAD_KCH, TUNE_P2P_SPLIT, TUNE_HAD_GROUPS, TUNE_HAD_TILES = 0, 1, 2, 3
LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib",
                        "libexafmm_t_amd.so")
TIMING_NAMES = ["P2M", "M2M", "P2L", "M2P", "P2P", "M2L", "L2L", "L2P", "FFT", "HAD", "IFFT"]

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class XfmmTree(ctypes.Structure):
    _fields_ = [
        ("nnodes", ctypes.c_int), ("depth", ctypes.c_int), ("r0", ctypes.c_double),
        ("x0", ctypes.c_double * 3),
        ("level_offset", _c_int_p), ("center", _c_dbl_p), ("child0", _c_int_p),
        ("src_begin", _c_int_p), ("src_count", _c_int_p), ("trg_begin", _c_int_p), ("trg_count", _c_int_p),
        ("nsrcs", ctypes.c_int), ("ntrgs", ctypes.c_int),
        ("src_coord", _c_dbl_p), ("trg_coord", _c_dbl_p),
        ("p2p_offset", _c_int_p), ("p2p_source", _c_int_p),
        ("m2p_offset", _c_int_p), ("m2p_source", _c_int_p),
        ("p2l_offset", _c_int_p), ("p2l_source", _c_int_p),
        ("m2l_source", _c_int_p),
    ]


_lib = None


def load(path=LIB_PATH):
    """Load the engine library (raises OSError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError("gfx950 engine library not built: %s (run exafmm-t_amd/build.py)" % path)
        lib = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        sig = {
            "xfmm_device_count": ([_c_int_p], ctypes.c_int),
            "xfmm_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
            "xfmm_destroy": ([vp], None),
            "xfmm_last_error": ([vp], ctypes.c_char_p),
            "xfmm_sizes": ([vp, _c_int_p, _c_int_p, _c_int_p], ctypes.c_int),
            "xfmm_set_operators": ([vp, ctypes.c_int, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp], ctypes.c_int),
            "xfmm_set_tree": ([vp, ctypes.POINTER(XfmmTree)], ctypes.c_int),
            "xfmm_set_charges_host": ([vp, vp], ctypes.c_int),
            "xfmm_set_charges_device": ([vp, vp], ctypes.c_int),
            "xfmm_clear_values": ([vp], ctypes.c_int),
            "xfmm_upward_pass": ([vp], ctypes.c_int),
            "xfmm_downward_pass": ([vp], ctypes.c_int),
            "xfmm_graph_capture": ([vp], ctypes.c_int),
            "xfmm_graph_launch": ([vp], ctypes.c_int),
            "xfmm_get_values": ([vp, vp], ctypes.c_int),
            "xfmm_get_equiv": ([vp, vp, vp], ctypes.c_int),
            "xfmm_set_equiv": ([vp, vp, vp], ctypes.c_int),
            "xfmm_values_device": ([vp, ctypes.POINTER(vp)], ctypes.c_int),
            "xfmm_synchronize": ([vp], ctypes.c_int),
            "xfmm_timings": ([vp, vp, ctypes.c_int], ctypes.c_int),
            "xfmm_set_value_scale": ([vp, vp], ctypes.c_int),
            "xfmm_set_tuning": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "xfmm_get_tuning": ([vp, ctypes.c_int, _c_int_p], ctypes.c_int),
            "xfmm_autotune": ([vp, ctypes.c_int, vp], ctypes.c_int),
            "xfmm_direct": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, ctypes.c_int,
                             vp, ctypes.c_int, vp, ctypes.c_int], ctypes.c_int),
        }
        for op in ("P2M", "M2M", "P2L", "M2P", "P2P", "M2L", "L2L", "L2P"):
            sig["xfmm_" + op] = ([vp], ctypes.c_int)
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


class XfmmError(RuntimeError):
    pass


def device_count():
    """HIP devices visible to the engine (0 without a GPU or without the library)."""
    try:
        lib = load()
    except OSError:
        return 0
    n = ctypes.c_int(0)
    lib.xfmm_device_count(ctypes.byref(n))
    return n.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _vals(a, cplx):
    """Values of kind T as a contiguous float64 buffer (complex interleaved)."""
    if cplx:
        return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)
    return np.ascontiguousarray(a, dtype=np.float64)


def direct(kernel, src, q, trg, wavek=0.0, precision=FP64, gradient=True, out=None):
    """Device direct sum (xfmm_direct): returns (nt, 4) or (nt,) values."""
    lib = load()
    cplx = kernel == HELMHOLTZ
    wk = complex(wavek)
    src = _f64(src).reshape(-1, 3)
    trg = _f64(trg).reshape(-1, 3)
    nt = trg.shape[0]
    w = 4 if gradient else 1
    dt = np.complex128 if cplx else np.float64
    if out is None:
        out = np.zeros((nt, w) if gradient else (nt,), dtype=dt)
    qv = _vals(q, cplx)
    ov = out.view(np.float64)
    rc = lib.xfmm_direct(kernel, precision, wk.real, wk.imag, _ptr(src), _ptr(qv), src.shape[0], _ptr(trg), nt,
                         _ptr(ov), 1 if gradient else 0)
    if rc != 0:
        raise XfmmError("xfmm_direct failed (%d): %s" % (rc, lib.xfmm_last_error(None).decode()))
    return out


class Engine:
    """One xfmm_context: operators, tree and device buffers of one FMM."""

    def __init__(self, kernel, p, wavek=0.0, precision=FP64, device=0):
        self.lib = load()
        self.kernel, self.p, self.precision = kernel, p, precision
        self.cplx = kernel == HELMHOLTZ
        wk = complex(wavek)
        h = ctypes.c_void_p()
        rc = self.lib.xfmm_create(kernel, p, wk.real, wk.imag, precision, device, ctypes.byref(h))
        if rc != 0:
            raise XfmmError("xfmm_create failed (%d): %s" % (rc, self.lib.xfmm_last_error(None).decode()))
        self.h = h
        ns, nc, nf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.lib.xfmm_sizes(h, ctypes.byref(ns), ctypes.byref(nc), ctypes.byref(nf))
        self.nsurf, self.nconv, self.nfreq = ns.value, nc.value, nf.value
        self._keep = []
        self.nnodes = self.nsrcs = self.ntrgs = 0

    def _check(self, rc, what):
        if rc != 0:
            raise XfmmError("%s failed (%d): %s" % (what, rc, self.lib.xfmm_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.xfmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_operators(self, ops):
        arrs = [_vals(ops[k], self.cplx) for k in ("uc2e_u", "uc2e_v", "dc2e_u", "dc2e_v", "m2m", "l2l")]
        m2l = _f64(ops["m2l"])
        self._check(self.lib.xfmm_set_operators(self.h, int(ops["nlevels"]), *[_ptr(a) for a in arrs],
                                                int(ops["nm2l"]), _ptr(m2l)), "xfmm_set_operators")

    def set_tree(self, f):
        keep = {k: (_i32(v) if k not in ("center", "src_coord", "trg_coord") else _f64(v))
                for k, v in f.items() if isinstance(v, np.ndarray)}
        t = XfmmTree()
        t.nnodes, t.depth, t.r0 = int(f["nnodes"]), int(f["depth"]), float(f["r0"])
        for d in range(3):
            t.x0[d] = float(f["x0"][d])
        for k in ("level_offset", "child0", "src_begin", "src_count", "trg_begin", "trg_count", "p2p_offset",
                  "p2p_source", "m2p_offset", "m2p_source", "p2l_offset", "p2l_source", "m2l_source"):
            setattr(t, k, keep[k].ctypes.data_as(_c_int_p))
        for k in ("center", "src_coord", "trg_coord"):
            setattr(t, k, keep[k].ctypes.data_as(_c_dbl_p))
        t.nsrcs = keep["src_coord"].size // 3
        t.ntrgs = keep["trg_coord"].size // 3
        self._check(self.lib.xfmm_set_tree(self.h, ctypes.byref(t)), "xfmm_set_tree")
        self._keep = [keep, t]
        self.nnodes, self.nsrcs, self.ntrgs = t.nnodes, t.nsrcs, t.ntrgs

    def set_charges(self, q):
        qv = _vals(q, self.cplx)
        self._check(self.lib.xfmm_set_charges_host(self.h, _ptr(qv)), "xfmm_set_charges_host")

    def set_charges_device(self, ptr):
        self._check(self.lib.xfmm_set_charges_device(self.h, ctypes.c_void_p(ptr)), "xfmm_set_charges_device")

    def upward_pass(self):
        self._check(self.lib.xfmm_upward_pass(self.h), "xfmm_upward_pass")

    def downward_pass(self):
        self._check(self.lib.xfmm_downward_pass(self.h), "xfmm_downward_pass")

    def graph_capture(self):
        self._check(self.lib.xfmm_graph_capture(self.h), "xfmm_graph_capture")

    def graph_launch(self):
        self._check(self.lib.xfmm_graph_launch(self.h), "xfmm_graph_launch")

    def op(self, name):
        self._check(getattr(self.lib, "xfmm_" + name)(self.h), "xfmm_" + name)

    def clear(self):
        self._check(self.lib.xfmm_clear_values(self.h), "xfmm_clear_values")

    def synchronize(self):
        self._check(self.lib.xfmm_synchronize(self.h), "xfmm_synchronize")

    def values(self):
        dt = np.complex128 if self.cplx else np.float64
        out = np.zeros((self.ntrgs, 4), dtype=dt)
        self._check(self.lib.xfmm_get_values(self.h, _ptr(out.view(np.float64))), "xfmm_get_values")
        return out

    def equiv(self):
        dt = np.complex128 if self.cplx else np.float64
        up = np.zeros((self.nnodes, self.nsurf), dtype=dt)
        dn = np.zeros((self.nnodes, self.nsurf), dtype=dt)
        self._check(self.lib.xfmm_get_equiv(self.h, _ptr(up.view(np.float64)), _ptr(dn.view(np.float64))),
                    "xfmm_get_equiv")
        return up, dn

    def set_equiv(self, up=None, dn=None):
        u = _vals(up, self.cplx) if up is not None else None
        d = _vals(dn, self.cplx) if dn is not None else None
        self._check(self.lib.xfmm_set_equiv(self.h, _ptr(u), _ptr(d)), "xfmm_set_equiv")

    def values_device_ptr(self):
        p = ctypes.c_void_p()
        self._check(self.lib.xfmm_values_device(self.h, ctypes.byref(p)), "xfmm_values_device")
        return p.value

    def set_value_scale(self, scale):
        """Per-target (potential, gradient) factors [ntrgs][2] applied after each evaluation (None removes)."""
        if scale is None:
            self._check(self.lib.xfmm_set_value_scale(self.h, None), "xfmm_set_value_scale")
            return
        sc = _f64(scale).reshape(-1)
        if sc.size != 2 * self.ntrgs:
            raise ValueError("value scale needs 2 factors per target")
        self._check(self.lib.xfmm_set_value_scale(self.h, _ptr(sc)), "xfmm_set_value_scale")

    def set_tuning(self, param, value):
        self._check(self.lib.xfmm_set_tuning(self.h, int(param), int(value)), "xfmm_set_tuning")

    def get_tuning(self, param):
        v = ctypes.c_int()
        self._check(self.lib.xfmm_get_tuning(self.h, int(param), ctypes.byref(v)), "xfmm_get_tuning")
        return v.value

    def autotune(self, reps=3):
        """Runs the engine's autotuner; returns {param: chosen value} and the best times (ms)."""
        best = np.zeros(2)
        self._check(self.lib.xfmm_autotune(self.h, int(reps), _ptr(best)), "xfmm_autotune")
        return ({"had_kch": self.get_tuning(TUNE_HAD_KCH), "p2p_split": self.get_tuning(TUNE_P2P_SPLIT),
                 "had_groups": self.get_tuning(TUNE_HAD_GROUPS), "had_tiles": self.get_tuning(TUNE_HAD_TILES)},
                {"p2p_l2p_ms": float(best[0]), "hadamard_ms": float(best[1])})

    def timings(self):
        ms = np.zeros(len(TIMING_NAMES))
        n = self.lib.xfmm_timings(self.h, _ptr(ms), len(ms))
        return dict(zip(TIMING_NAMES[:n], ms[:n].tolist()))
