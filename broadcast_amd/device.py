"""Device plumbing: contexts, streams and array staging.

PyTorch-ROCm provides device memory and the current HIP stream; the compute is
the HIP library behind the C ABI.  Arrays cross the boundary in the
reference's layout (Fortran order, ghost frame included):

* a torch tensor on a HIP device (float64 / int32) whose memory is
  Fortran-contiguous (or a 1-D tensor) is passed by pointer, zero copy;
* a numpy array (the f2py call surface) is staged to the device, the kernel
  runs, and outputs are copied back into the same numpy buffer in place, so a
  caller sees the f2py ``intent(inout)`` semantics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_CTX = {}


def torch():
    import torch as _t

    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("broadcast_amd needs a HIP device (MI355X / gfx950); none is visible")
    return t


def current_stream_ptr():
    t = torch()
    return C.c_void_p(t.cuda.current_stream().cuda_stream)


class Context:
    """One bca_ctx per (im, jm, gh, device); kept for the life of the process."""

    def __init__(self, im, jm, gh):
        require_gpu()
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.bca_ctx_create(C.byref(h), im, jm, gh, current_stream_ptr()), "bca_ctx_create")
        self.h = h
        self.im, self.jm, self.gh = im, jm, gh

    def sync_stream(self):
        _lib.check(_lib.lib().bca_ctx_set_stream(self.h, current_stream_ptr()), "bca_ctx_set_stream")
        return self.h

    @property
    def scratch_bytes(self):
        return int(_lib.lib().bca_ctx_scratch_bytes(self.h))


def context(im, jm, gh) -> Context:
    t = require_gpu()
    key = (int(im), int(jm), int(gh), t.cuda.current_device())
    c = _CTX.get(key)
    if c is None:
        c = Context(int(im), int(jm), int(gh))
        _CTX[key] = c
    c.sync_stream()
    return c


def _is_tensor(a):
    try:
        t = torch()
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, t.Tensor)


def _fortran_contiguous(t):
    if t.dim() <= 1:
        return t.is_contiguous()
    expect = 1
    for size, stride in zip(t.shape, t.stride()):
        if size != 1 and stride != expect:
            return False
        expect *= size
    return True


class Staged:
    """Device view of one argument; ``finish()`` writes numpy outputs back."""

    def __init__(self, a, dtype=np.float64, writeback=False):
        self.src = a
        self.writeback = writeback
        t = torch()
        tdt = t.float64 if dtype == np.float64 else t.int32
        if a is None:
            self.dev = None
        elif _is_tensor(a):
            if not a.is_cuda:
                raise TypeError("torch tensors must live on the HIP device")
            if a.dtype != tdt:
                raise TypeError(f"expected {tdt}, got {a.dtype}")
            if not _fortran_contiguous(a):
                raise TypeError("tensor memory must be Fortran-contiguous (the reference's layout)")
            self.dev = a
        else:
            arr = np.asarray(a)
            if arr.dtype != dtype:
                raise TypeError(f"expected {np.dtype(dtype)}, got {arr.dtype}")
            if not arr.flags.f_contiguous:
                raise TypeError("numpy arrays must be Fortran-ordered (order='F')")
            flat = arr.reshape(-1, order="F")
            self.dev = t.from_numpy(flat).to(device="cuda", non_blocking=False)

    @property
    def ptr(self):
        return None if self.dev is None else C.c_void_p(self.dev.data_ptr())

    def finish(self):
        if self.writeback and self.dev is not None and not _is_tensor(self.src):
            flat = np.asarray(self.src).reshape(-1, order="F")
            flat[...] = self.dev.cpu().numpy()


def stage(*args, out=()):
    """Stage a list of arrays; indices listed in ``out`` are written back."""
    return [Staged(a, np.int32 if (isinstance(a, np.ndarray) and a.dtype == np.int32) or
                   (_is_tensor(a) and str(a.dtype) == "torch.int32") else np.float64,
                   writeback=(k in out)) for k, a in enumerate(args)]


def finish(staged):
    t = torch()
    t.cuda.current_stream().synchronize()
    for s in staged:
        s.finish()


def int4(box):
    """Interface box (2,2) [[imin, jmin], [imax, jmax]] (float64 as the driver
    builds it, or int) -> C int[4] {imin, jmin, imax, jmax}."""
    b = np.asarray(box)
    return (C.c_int * 4)(int(b[0, 0]), int(b[0, 1]), int(b[1, 0]), int(b[1, 1]))


LOC = {"Ilo": 0, "Ihi": 1, "Jlo": 2, "Jhi": 3}


def loc_code(loc):
    if isinstance(loc, bytes):
        loc = loc.decode()
    loc = str(loc).strip()
    if loc not in LOC:
        raise ValueError(f"loc must be one of {sorted(LOC)}, got {loc!r}")
    return LOC[loc]


def phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4):
    return _lib.Phys(float(cp), float(cv), float(prandtl), float(gam), float(rgaz), float(cs),
                     float(muref), float(tref), float(s_suth), float(k2), float(k4))
