"""f_dz: spanwise operators of the 3D linearisation (srcfv/dz,
srcfv/compile_dz.py).  Call forms of BROADCAST_npz.py:1242-1243."""
from broadcast_amd import _lib
from broadcast_amd import device as dev


def _dz(which, out, w, wd, nx, ny, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, im, jm):
    ctx = dev.context(im, jm, gh)
    st = dev.stage(out, w, wd, nx, ny, vol, volf, out=(0,))
    g = _lib.Metrics(st[3].ptr, st[4].ptr, st[5].ptr, st[6].ptr)
    ph = dev.phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, 0.0, 0.0)
    dz, dz2 = (st[0].ptr, None) if which == 1 else (None, st[0].ptr)
    _lib.check(_lib.lib().bca_dz_5p(ctx.h, dz, dz2, st[1].ptr, st[2].ptr, g, ph),
               "coeffs_5p_dz" if which == 1 else "coeffs_5p_dz2")
    dev.finish(st)


def coeffs_5p_dz(dz, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs, muref, tref,
                 s_suth, im, jm):
    """srcfv/dz/coeffs_5p_dz.F90:10-174: dz on the interior"""
    _dz(1, dz, w, wd, nx, ny, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, im, jm)


def coeffs_5p_dz2(dz2, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs, muref, tref,
                  s_suth, im, jm):
    """srcfv/dz/coeffs_5p_dz2.F90: dz2 on the interior"""
    _dz(2, dz2, w, wd, nx, ny, vol, volf, gh, cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, im, jm)
