"""f_sch: the DNC residual (srcfv/rhs, srcfv/compile_rhs.py:42-44)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev


def flux_num_dnc5_2d(res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs,
                     muref, tref, s_suth, k2, k4, im, jm):
    """srcfv/rhs/flux_num_dnc5.F90:7-226; residu written on the interior only.
    x0, y0, xc, yc are accepted for the call form but, as in the reference,
    never read."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(res, w, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[2].ptr, st[3].ptr, st[4].ptr, st[5].ptr)
    ph = dev.phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
    _lib.check(_lib.lib().bca_residual_dnc5(ctx.h, st[0].ptr, st[1].ptr, g, ph), "flux_num_dnc5_2d")
    dev.finish(st)
