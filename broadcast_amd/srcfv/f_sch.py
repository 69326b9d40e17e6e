"""f_sch: the DNC residuals (srcfv/rhs, srcfv/compile_rhs.py:42-44)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev


def _residual(entry, name, res, w, nx, ny, vol, volf, gh, phys, im, jm):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(res, w, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[2].ptr, st[3].ptr, st[4].ptr, st[5].ptr)
    ph = dev.phys(*phys)
    _lib.check(getattr(_lib.lib(), entry)(ctx.h, st[0].ptr, st[1].ptr, g, ph), name)
    dev.finish(st)


def flux_num_dnc5_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                     muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc5.F90:7-226; residu written on the interior only.
    x0, y0, xc, yc are accepted for the call form but, as in the reference,
    never read."""
    _residual("bca_residual_dnc5", "flux_num_dnc5_2d", res, w, nx, ny, vol, volf, gh,
              (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc5_nowall_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                            muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc5_nowall.F90:7-176: interior formulas on every
    face, no wall rows; residu written on the interior only."""
    _residual("bca_residual_dnc5_nowall", "flux_num_dnc5_nowall_2d", res, w, nx, ny, vol, volf, gh,
              (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc3_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                     muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc3.F90:7-219, on arrays of ghost width gh = 2."""
    _residual("bca_residual_dnc3", "flux_num_dnc3_2d", res, w, nx, ny, vol, volf, gh,
              (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc3_nowall_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                            muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc3_nowall.F90:7, on arrays of ghost width gh = 2."""
    _residual("bca_residual_dnc3_nowall", "flux_num_dnc3_nowall_2d", res, w, nx, ny, vol, volf, gh,
              (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def _residual_order(order, nowall, name, res, w, nx, ny, vol, volf, gh, phys, im, jm):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(res, w, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[2].ptr, st[3].ptr, st[4].ptr, st[5].ptr)
    ph = dev.phys(*phys)
    _lib.check(_lib.lib().bca_residual_dnc(ctx.h, order, int(nowall), st[0].ptr, st[1].ptr, g, ph), name)
    dev.finish(st)


def flux_num_dnc7_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                     muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc7.F90, on arrays of ghost width gh = 4."""
    _residual_order(7, False, "flux_num_dnc7_2d", res, w, nx, ny, vol, volf, gh,
                    (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc7_nowall_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                            muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc7_nowall.F90, on arrays of ghost width gh = 4."""
    _residual_order(7, True, "flux_num_dnc7_nowall_2d", res, w, nx, ny, vol, volf, gh,
                    (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc9_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                     muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc9.F90, on arrays of ghost width gh = 5."""
    _residual_order(9, False, "flux_num_dnc9_2d", res, w, nx, ny, vol, volf, gh,
                    (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)


def flux_num_dnc9_nowall_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                            muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc9_nowall.F90, on arrays of ghost width gh = 5."""
    _residual_order(9, True, "flux_num_dnc9_nowall_2d", res, w, nx, ny, vol, volf, gh,
                    (cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4), im, jm)
