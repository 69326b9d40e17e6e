"""f_geom: mesh metrics (srcfv/geom, srcfv/compile_geom.py:32)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev


def computegeom_2d(x0, y0, nx, ny, xc, yc, vol, volf, im, jm, gh):
    """srcfv/geom/computegeom.F90:3-104 (all arrays in/out)."""
    ctx = dev.context(im, jm, gh)
    st = dev.stage(x0, y0, nx, ny, xc, yc, vol, volf, out=tuple(range(8)))
    _lib.check(_lib.lib().bca_geom(ctx.h, *[s.ptr for s in st]), "computegeom_2d")
    dev.finish(st)
