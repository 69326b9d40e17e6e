"""f_adj: adjoint routines (srcfv/adjoint, srcfv/compile_adjoint.py).

Same names and call forms as the TAPENADE-generated _b routines: the flux
adjoint the cylinder driver resolves (cylinder.py:1141) and the boundary
condition adjoints of its adjoint Jacobian loop (cylinder.py:1121-1154).
Each BC adjoint pushes the adjoint of the ghost cells it writes to the cells
it reads (wb += (d ghosts / d w)^T wb(ghosts)) and zeroes wb on its ghosts;
w is left as given (the reference restores it on exit)."""
import numpy as np

from broadcast_amd import _lib
from broadcast_amd import device as dev
from broadcast_amd.srcfv import _bc


def flux_num_dnc5_2d_b(res, resb, w, wb, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/adjoint/flux_num_dnc5_b.f90:15-6676: wb = (dR/dw)^T resb over the
    whole array (ghost frame included, overwritten); resb's interior is
    consumed (zeroed) and res is left untouched, as in the reference."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(resb, wb, w, nx, ny, vol, volf, out=(0, 1))
    g = _lib.Metrics(st[3].ptr, st[4].ptr, st[5].ptr, st[6].ptr)
    ph = dev.phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
    _lib.check(_lib.lib().bca_vjp_dnc5(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, g, ph), "flux_num_dnc5_2d_b")
    dev.finish(st)


def bc_supandsubinlet_2d_b(w, wb, loc, interf, field, nx, ny, gam, im, jm, lm=None, gh=None):
    """srcfv/adjoint/bc_supandsubinlet_b.f90:11"""
    gh = gh if gh is not None else field.shape[1]
    lm = lm if lm is not None else field.shape[0]
    _bc.adjoint(_lib.BC_INLET, w, wb, im, jm, gh, loc, interf, data=field, lm=lm, gam=gam, nx=nx, ny=ny,
                what="bc_supandsubinlet_2d_b")


def bc_no_reflexion_2d_b(w, wb, wbd, loc, interf, nx, ny, gam, gh, im, jm, lm=None):
    """srcfv/adjoint/bc_no_reflexion_b.f90:16"""
    lm = lm if lm is not None else wbd.shape[0]
    _bc.adjoint(_lib.BC_NOREF, w, wb, im, jm, gh, loc, interf, data=wbd, lm=lm, gam=gam, nx=nx, ny=ny,
                what="bc_no_reflexion_2d_b")


def bc_extrapolate_o2_2d_b(w, wb, loc, interf, im, jm, gh, em=None):
    """srcfv/adjoint/bc_extrapolateo2_b.f90:18 (5-variable state arrays)"""
    em = em if em is not None else (w.shape[2] if len(w.shape) == 3 else 1)
    if int(em) != 5:
        raise NotImplementedError("bc_extrapolate_o2_2d_b: 5-variable state arrays only")
    _bc.adjoint(_lib.BC_EXTRAP2, w, wb, im, jm, gh, loc, interf, what="bc_extrapolate_o2_2d_b")


def bc_wall_viscous_adia_2d_b(w, wb, loc, gam, interf, gh, im, jm):
    """srcfv/adjoint/bc_wall_viscous_b.f90:11"""
    _bc.adjoint(_lib.BC_WALL_ADIA, w, wb, im, jm, gh, loc, interf, gam=gam, what="bc_wall_viscous_adia_2d_b")


def _same(a, b):
    return a is b or (hasattr(a, "data_ptr") and hasattr(b, "data_ptr") and a.data_ptr() == b.data_ptr())


def jn_match_2d_b(wr, wrb, prr, gh1r, gh2r, gh3r, gh4r, imr, jmr, wd, wdb, prd, gh1d, gh2d, gh3d, gh4d,
                  imd, jmd, tr, em=None):
    """srcfv/adjoint/jn_match_2d_b.f90:11 for the in-block join (receiver and
    donor the same array, as every join of the O-mesh): the transpose of the
    copy, wb(prd) += wb(prr) then wb(prr) = 0.  (The reference routine zeroes
    the donor adjoint array before accumulating, which for one array would
    clear wb; its adjoint loop never calls it, cylinder.py:1121-1154.)"""
    if not (_same(wr, wd) and _same(wrb, wdb)):
        raise NotImplementedError("jn_match_2d_b: only the single-block join (receiver == donor) is supported")
    if len({gh1r, gh2r, gh3r, gh4r, gh1d, gh2d, gh3d, gh4d}) != 1 or (imr, jmr) != (imd, jmd):
        raise NotImplementedError("jn_match_2d_b: uniform ghost width and one block only")
    em = em if em is not None else (wr.shape[2] if len(wr.shape) == 3 else 1)
    if int(em) != 5 or list(np.asarray(tr).reshape(-1)) != [1, 2]:
        raise NotImplementedError("jn_match_2d_b: 5-variable state arrays, identity transform only")
    _bc.adjoint(_lib.BC_JOIN, wr, wrb, imr, jmr, int(gh1r), interf=prr, interf2=prd, what="jn_match_2d_b")
