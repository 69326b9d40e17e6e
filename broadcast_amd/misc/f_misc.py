"""f_misc: seeds and JVP -> COO scatter of the colored Jacobian
(misc/ComputeJacobian.f90, misc/compile_misc.py:36).  Call forms of
BROADCAST_npz.py:1072,1127,1245 and cylinder.py:916,978,1159."""
from __future__ import annotations

from broadcast_amd import _lib
from broadcast_amd import device as dev


def testvector(wd, m, l, k, gh, im, jm):
    """misc/ComputeJacobian.f90:357-374"""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(wd, out=(0,))
    _lib.check(_lib.lib().bca_seed(ctx.h, st[0].ptr, int(m), int(l), int(k)), "testvector")
    dev.finish(st)


def testvector_partial(wd, m, l, k, gh, im, jm, istart, iend, jstart, jend):
    """misc/ComputeJacobian.f90:1075-1092 (0-based index box)"""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(wd, out=(0,))
    _lib.check(_lib.lib().bca_seed_partial(ctx.h, st[0].ptr, int(m), int(l), int(k), int(istart), int(iend),
                                           int(jstart), int(jend)), "testvector_partial")
    dev.finish(st)


def computejacobianfromjv_relaxed_withjnandcheck(jac, ia, ja, resd, m, l, k, gh, coefdiag, mini, n, im=None,
                                                 jm=None, nbentry=None):
    """misc/ComputeJacobian.f90:1095-1204, call form of cylinder.py:939 (im, jm
    inferred from coefdiag)"""
    im = im if im is not None else coefdiag.shape[0]
    jm = jm if jm is not None else coefdiag.shape[1]
    ctx = dev.context(im, jm, gh)
    st = dev.stage(jac, ia, ja, resd, coefdiag, out=(0, 1, 2))
    nb = jac.shape[0] if hasattr(jac, "shape") else len(jac)
    _lib.check(_lib.lib().bca_scatter_jv_2zones(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, st[3].ptr, int(m), int(l),
                                                int(k), st[4].ptr, float(mini), int(n), int(nb)),
               "computejacobianfromjv_relaxed_withjnandcheck")
    dev.finish(st)


def _scatter(jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag, withjn):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(jac, ia, ja, resd, coefdiag, out=(0, 1, 2))
    nb = jac.shape[0] if hasattr(jac, "shape") else len(jac)
    _lib.check(_lib.lib().bca_scatter_jv(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, st[3].ptr, int(m),
                                         int(l), int(k), st[4].ptr, int(withjn), int(nb)),
               "computejacobianfromjv")
    dev.finish(st)


def computejacobianfromjv(jac, ia, ja, resd, m, l, k, gh, im, jm, nbentry=None):
    """misc/ComputeJacobian.f90:292-355"""
    _scatter(jac, ia, ja, resd, m, l, k, gh, im, jm, None, False)


def computejacobianfromjv_relaxed(jac, ia, ja, resd, m, l, k, gh, coefdiag, im=None, jm=None,
                                  nbentry=None):
    """misc/ComputeJacobian.f90:503-570 (im, jm inferred from coefdiag, as f2py does)"""
    im = im if im is not None else coefdiag.shape[0]
    jm = jm if jm is not None else coefdiag.shape[1]
    _scatter(jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag, False)


def computejacobianfromjv_relaxed_withjn(jac, ia, ja, resd, m, l, k, gh, coefdiag, im=None, jm=None,
                                         nbentry=None):
    """misc/ComputeJacobian.f90:847-926"""
    im = im if im is not None else coefdiag.shape[0]
    jm = jm if jm is not None else coefdiag.shape[1]
    _scatter(jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag, True)


def computejacobianfromjv_withjn(jac, ia, ja, resd, m, l, k, gh, im, jm, nbentry=None):
    """misc/ComputeJacobian.f90 (join-aware, not relaxed), cylinder.py:1159"""
    _scatter(jac, ia, ja, resd, m, l, k, gh, im, jm, None, True)


def computejacobianfromdz(jac, ia, ja, dz, m, l, k, gh, im, jm, nbentry=None):
    """misc/ComputeJacobian.f90:708-779 (call form of BROADCAST_npz.py:1245)"""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(jac, ia, ja, dz, out=(0, 1, 2))
    nb = jac.shape[0] if hasattr(jac, "shape") else len(jac)
    _lib.check(_lib.lib().bca_scatter_dz(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, st[3].ptr, int(m), int(l), int(k),
                                         int(nb)), "computejacobianfromdz")
    dev.finish(st)
