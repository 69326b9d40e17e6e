"""f_norm: residual norms (srcfv/norm.F90, srcfv/compile_norm.py:15)."""
import ctypes as C

import numpy as np

from broadcast_amd import _lib
from broadcast_amd import device as dev


def _norms(fn, res, im, jm, gh):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(res)
    out = (C.c_double * 10)()
    _lib.check(fn(ctx.h, st[0].ptr, out), "compute_norml2")
    o = np.array(out[:], dtype=np.float64)
    return o[:5].copy(), o[5:].copy()


def compute_norml2inf(res, im, jm, gh):
    """srcfv/norm.F90:34-77 -> (norm, ninf): per-variable L2 and (sum r^10)^0.1."""
    return _norms(_lib.lib().bca_norml2inf, res, im, jm, gh)


def compute_norml2(res, im, jm, gh):
    """srcfv/norm.F90:2-32 -> (norm, nmoy)."""
    return _norms(_lib.lib().bca_norml2, res, im, jm, gh)
