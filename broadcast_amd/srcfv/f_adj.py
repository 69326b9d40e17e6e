"""f_adj: adjoint routines (srcfv/adjoint, srcfv/compile_adjoint.py).

Same name and call form as the TAPENADE-generated _b routine the cylinder
driver resolves (cylinder.py:1141)."""
from .. import _lib
from .. import device as dev


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
