"""f_lin: tangent-linear routines (srcfv/tangent, srcfv/compile_tangent.py:5,32).

Same names and call forms as the TAPENADE-generated _d routines the reference
driver resolves by name (BROADCAST_npz.py:989-1001, 1079-1085)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev
from broadcast_amd.srcfv import _bc


def _tangent(entry, name, res, resd, w, wd, nx, ny, vol, volf, gh, phys, im, jm):
    """resd = dR/dw . wd on the interior.  res is left as given: the generated
    tangents assign only residud (their primal residual is never stored)."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(resd, w, wd, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[3].ptr, st[4].ptr, st[5].ptr, st[6].ptr)
    ph = dev.phys(*phys)
    _lib.check(getattr(_lib.lib(), entry)(ctx.h, None, st[0].ptr, st[1].ptr, st[2].ptr, g, ph), name)
    dev.finish(st)


def flux_num_dnc5_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc5_d.f90:15-3870"""
    _tangent("bca_jvp_dnc5", "flux_num_dnc5_2d_d", res, resd, w, wd, nx, ny, vol, volf, gh,
             (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc5_nowall_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl,
                              gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc5_nowall_d.f90:15"""
    _tangent("bca_jvp_dnc5_nowall", "flux_num_dnc5_nowall_2d_d", res, resd, w, wd, nx, ny, vol, volf, gh,
             (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc3_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc3_d.f90:15, on arrays of ghost width gh = 2"""
    _tangent("bca_jvp_dnc3", "flux_num_dnc3_2d_d", res, resd, w, wd, nx, ny, vol, volf, gh,
             (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc3_nowall_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl,
                              gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc3_nowall_d.f90:15, on arrays of ghost width gh = 2"""
    _tangent("bca_jvp_dnc3_nowall", "flux_num_dnc3_nowall_2d_d", res, resd, w, wd, nx, ny, vol, volf, gh,
             (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def _tangent_order(order, nowall, name, resd, w, wd, nx, ny, vol, volf, gh, phys, im, jm):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(resd, w, wd, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[3].ptr, st[4].ptr, st[5].ptr, st[6].ptr)
    ph = dev.phys(*phys)
    _lib.check(_lib.lib().bca_jvp_dnc(ctx.h, order, int(nowall), None, st[0].ptr, st[1].ptr, st[2].ptr, g, ph), name)
    dev.finish(st)


def flux_num_dnc7_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc7_d.f90, on arrays of ghost width gh = 4"""
    _tangent_order(7, False, "flux_num_dnc7_2d_d", resd, w, wd, nx, ny, vol, volf, gh,
                   (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc7_nowall_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl,
                              gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc7_nowall_d.f90, on arrays of ghost width gh = 4"""
    _tangent_order(7, True, "flux_num_dnc7_nowall_2d_d", resd, w, wd, nx, ny, vol, volf, gh,
                   (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc9_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam,
                       rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc9_d.f90, on arrays of ghost width gh = 5"""
    _tangent_order(9, False, "flux_num_dnc9_2d_d", resd, w, wd, nx, ny, vol, volf, gh,
                   (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc9_nowall_2d_d(res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl,
                              gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/tangent/flux_num_dnc9_nowall_d.f90, on arrays of ghost width gh = 5"""
    _tangent_order(9, True, "flux_num_dnc9_nowall_2d_d", resd, w, wd, nx, ny, vol, volf, gh,
                   (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


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
