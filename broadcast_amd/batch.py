"""Residual + JVP of a batch of meshes of different sizes, launched together.

Per mesh the work is the reference's per-iteration sequence -- the BC calls in
order, then flux_num_dnc5_2d (BROADCAST_npz.py:1018-1031) or, with a tangent,
the linearised BCs and flux_num_dnc5_2d_d (:1079-1085) -- as ``Flow.residual``
/ ``Flow.jvp`` do one mesh at a time; ``UnevenBatch`` runs all meshes
together, one launch per stage over the concatenated work of every mesh
(bca_batch_residual_jvp), so small meshes do not each pay a chain of
launches on a mostly idle chip.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from . import device as dev
from .flow import Flow, _ptr


class UnevenBatch:
    MAX = 16

    def __init__(self, flows):
        flows = list(flows)
        if not 1 <= len(flows) <= self.MAX:
            raise ValueError(f"1..{self.MAX} meshes per batch")
        if any(not isinstance(f, Flow) for f in flows):
            raise TypeError("UnevenBatch takes Flow objects")
        self.flows = flows
        self.L = _lib.lib()
        h = C.c_void_p()
        _lib.check(self.L.bca_batch_create(C.byref(h), dev.current_stream_ptr()), "bca_batch_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.bca_batch_destroy(h)
            self.h = None

    @property
    def cells(self):
        return sum(f.im * f.jm for f in self.flows)

    def run(self, ws, wds=None, res=None, resd=None):
        """BCs in place on every ws[k] (and wds[k]), then res[k] = R(ws[k]) and
        resd[k] = dR/dw . wds[k] on the interiors.  Returns (res, resd)."""
        n = len(self.flows)
        if len(ws) != n or (wds is not None and len(wds) != n):
            raise ValueError("one state (and tangent) per mesh")
        res = [f.new_state() for f in self.flows] if res is None else res
        if wds is not None and resd is None:
            resd = [f.new_state() for f in self.flows]
        items = (_lib.BatchItem * n)()
        for k, f in enumerate(self.flows):
            for a in (ws[k], res[k]) + ((wds[k], resd[k]) if wds is not None else ()):
                if a.numel() != f.ni * f.nj * 5 or a.dtype != f.t.float64 or not a.is_cuda or not a.is_contiguous():
                    raise ValueError(f"mesh {k}: arrays must be contiguous float64 device tensors of ni*nj*5")
            it = items[k]
            it.im, it.jm = f.im, f.jm
            it.g = f.metrics
            it.phys = f.phys
            it.ops = C.cast(f.ops, C.POINTER(_lib.BcOp))
            it.nops = f.nops
            it.w, it.residu = ws[k].data_ptr(), res[k].data_ptr()
            if wds is not None:
                it.wd, it.residud = wds[k].data_ptr(), resd[k].data_ptr()
        _lib.check(self.L.bca_batch_set_stream(self.h, dev.current_stream_ptr()), "bca_batch_set_stream")
        _lib.check(self.L.bca_batch_residual_jvp(self.h, items, n), "bca_batch_residual_jvp")
        return res, resd
