"""f_init: initialisation helpers (compilef90.py:17)."""
from broadcast_amd import _lib
from broadcast_amd import device as dev


def set_bndbl_2d(w, field, wbd, im, jm=None, gh=None):
    """set_bnd.f90:2-24: field(j,d,:) = w(1-d,j,:), wbd(i,:) = w(i-gh,jm,:)."""
    gh = gh if gh is not None else field.shape[1]
    jm = jm if jm is not None else field.shape[0]
    ctx = dev.context(im, jm, gh)
    st = dev.stage(w, field, wbd, out=(1, 2))
    _lib.check(_lib.lib().bca_set_bndbl(ctx.h, st[0].ptr, st[1].ptr, st[2].ptr), "set_bndbl_2d")
    dev.finish(st)
