"""f_lin: tangent-linear routines (srcfv/tangent, srcfv/compile_tangent.py:5,32).

Same names and call forms as the TAPENADE-generated _d routines the reference
driver resolves by name (BROADCAST_npz.py:989-1001, 1079-1085)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev
from broadcast_amd.srcfv import _bc


def flux_num_dnc5_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc5_d.f90:15-3870: res and resd = dR/dw . wd."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(res, resd, w, wd, nx, ny, vol, volf, out=(0, 1))
    g = _lib.Metrics(st[4].ptr, st[5].ptr, st[6].ptr, st[7].ptr)
    ph = dev.phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
    _lib.check(_lib.lib().bca_jvp_dnc5(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr, st[3].ptr, g, ph),
               "flux_num_dnc5_2d_d")
    dev.finish(st)


def bc_supandsubinlet_2d_d(w, wd0, loc, interf, field, nx, ny, gam, im, jm, lm=None, gh=None):
    """srcfv/tangent/bc_supandsubinlet_d.f90:11"""
    _bc.inlet(w, wd0, loc, interf, field, nx, ny, gam, im, jm, lm, gh)


def bc_no_reflexion_2d_d(w, wd, wbd, loc, interf, nx, ny, gam, gh, im, jm, lm=None):
    """srcfv/tangent/bc_no_reflexion_d.f90:16"""
    _bc.noref(w, wd, wbd, loc, interf, nx, ny, gam, gh, im, jm, lm)


def bc_extrapolate_o2_2d_d(w, wd, loc, interf, im, jm, gh, em=None):
    """srcfv/tangent/bc_extrapolateo2_d.f90:18"""
    _bc.extrap2(w, wd, loc, interf, im, jm, gh, em)


def bc_wall_viscous_adia_2d_d(w, wd, loc, gam, interf, gh, im, jm):
    """srcfv/tangent/bc_wall_viscous_d.f90:11"""
    _bc.wall(w, wd, loc, gam, interf, gh, im, jm)
