"""Shared implementation of the boundary-condition entry points of f_bnd
(primal) and f_lin (linearised, w and wd both updated)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from broadcast_amd import _lib
from broadcast_amd import device as dev


def inlet(w, wd, loc, interf, field, nx, ny, gam, im, jm, lm=None, gh=None):
    gh = gh if gh is not None else field.shape[1]
    lm = lm if lm is not None else field.shape[0]
    ctx = dev.context(im, jm, gh)
    out = (0, 1) if wd is not None else (0,)
    st = dev.stage(w, wd, field, nx, ny, out=out)
    _lib.check(_lib.lib().bca_bc_inlet(ctx.h, st[0].ptr, st[1].ptr, dev.loc_code(loc), dev.int4(interf),
                                       st[2].ptr, int(lm), st[3].ptr, st[4].ptr, float(gam)),
               "bc_supandsubinlet_2d")
    dev.finish(st)


def noref(w, wd, wbd, loc, interf, nx, ny, gam, gh, im, jm, lm=None):
    lm = lm if lm is not None else wbd.shape[0]
    ctx = dev.context(im, jm, gh)
    out = (0, 1) if wd is not None else (0,)
    st = dev.stage(w, wd, wbd, nx, ny, out=out)
    _lib.check(_lib.lib().bca_bc_noref(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, int(lm), dev.loc_code(loc),
                                       dev.int4(interf), st[3].ptr, st[4].ptr, float(gam)),
               "bc_no_reflexion_2d")
    dev.finish(st)


def extrap2(w, wd, loc, interf, im, jm, gh, em=None):
    em = em if em is not None else (w.shape[2] if len(w.shape) == 3 else 1)
    ctx = dev.context(im, jm, gh)
    out = (0, 1) if wd is not None else (0,)
    st = dev.stage(w, wd, out=out)
    _lib.check(_lib.lib().bca_bc_extrap2(ctx.h, st[0].ptr, st[1].ptr, dev.loc_code(loc),
                                         dev.int4(interf), int(em)), "bc_extrapolate_o2_2d")
    dev.finish(st)


def wall(w, wd, loc, gam, interf, gh, im, jm):
    ctx = dev.context(im, jm, gh)
    out = (0, 1) if wd is not None else (0,)
    st = dev.stage(w, wd, out=out)
    _lib.check(_lib.lib().bca_bc_wall_adia(ctx.h, st[0].ptr, st[1].ptr, dev.loc_code(loc),
                                           dev.int4(interf), float(gam)), "bc_wall_viscous_adia_2d")
    dev.finish(st)


def join(wr, prr, gh1r, gh2r, gh3r, gh4r, imr, jmr, wd, prd, gh1d, gh2d, gh3d, gh4d, imd, jmd, tr,
         em=None):
    if not (wr is wd or (hasattr(wr, "data_ptr") and hasattr(wd, "data_ptr")
                         and wr.data_ptr() == wd.data_ptr())):
        raise NotImplementedError("jn_match_2d: only the single-block join (receiver == donor) is supported")
    if len({gh1r, gh2r, gh3r, gh4r, gh1d, gh2d, gh3d, gh4d}) != 1 or (imr, jmr) != (imd, jmd):
        raise NotImplementedError("jn_match_2d: uniform ghost width and one block only")
    gh = int(gh1r)
    em = em if em is not None else (wr.shape[2] if len(wr.shape) == 3 else 1)
    # the array may be a cell array (imr = im) or a face array (imr = im + 1)
    ni_arr, nj_arr = int(imr) + 2 * gh, int(jmr) + 2 * gh
    ctx = dev.context(imr, jmr, gh)
    st = dev.stage(wr, out=(0,))
    tr_c = (C.c_int * 2)(int(np.asarray(tr)[0]), int(np.asarray(tr)[1]))
    _lib.check(_lib.lib().bca_bc_join(ctx.h, st[0].ptr, dev.int4(prr), dev.int4(prd), tr_c, int(em),
                                      ni_arr, nj_arr), "jn_match_2d")
    dev.finish(st)


# ------------------------------------------------------------------ adjoints
def _op(kind, loc=0, interf=None, interf2=None, data=None, lm=0, gam=0.0):
    op = _lib.BcOp()
    op.kind, op.loc, op.gam, op.lm = kind, loc, float(gam), int(lm)
    if interf is not None:
        op.box[:] = list(dev.int4(interf))
    if interf2 is not None:
        op.box2[:] = list(dev.int4(interf2))
    op.tr[:] = [1, 2]
    op.data = data
    return op


def adjoint(kind, w, wb, im, jm, gh, loc=None, interf=None, interf2=None, data=None, lm=0, gam=0.0,
            nx=None, ny=None, what="bc adjoint"):
    """wb updated in place: the BC's ghost adjoints pushed to the cells it
    reads, then zeroed (w is read only)."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(w, wb, data, nx, ny, out=(1,))
    dp = st[2].ptr
    op = _op(kind, dev.loc_code(loc) if loc is not None else 0, interf, interf2,
             None if dp is None else dp.value, lm, gam)
    _lib.check(_lib.lib().bca_bc_adjoint(ctx.h, st[0].ptr, st[1].ptr, C.byref(op), st[3].ptr, st[4].ptr), what)
    dev.finish(st)
