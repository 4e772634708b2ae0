"""exafmm: the exafmm-t Python API on the MI355X (gfx950) evaluator.

Submodules mirror the reference module (python/exafmm.cpp:325-537):
``exafmm.laplace``, ``exafmm.helmholtz``, ``exafmm.modified_helmholtz``.
Evaluation always runs on the GPU through the C ABI of
include/exafmm_t_amd.h; importing fails if the native module is missing.
"""
import os
import sys

_here = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _exafmm  # noqa: F401
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "exafmm native module is not built; run `python exafmm-t_amd/build.py` "
        "(needs hipcc for gfx950): %s" % e) from e

laplace = _exafmm.laplace
helmholtz = _exafmm.helmholtz
modified_helmholtz = _exafmm.modified_helmholtz
for _name in ("laplace", "helmholtz", "modified_helmholtz"):
    sys.modules[__name__ + "." + _name] = getattr(_exafmm, _name)

FP64 = 0
MIXED = 1
FLOAT = 2

__all__ = ["laplace", "helmholtz", "modified_helmholtz", "FP64", "MIXED", "FLOAT"]
