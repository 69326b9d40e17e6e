"""The per-iteration Jacobian stage captured in a HIP graph.

BCs + residual + colored Jacobian into the reference's COO
(BROADCAST_npz.py:1018-1127, cylinder.py:853-985) is ~40 kernel launches with
no host synchronisation, so it is captured once (torch.cuda.graph on the
stream the library launches on) and replayed every iteration; the state w,
the relaxation coefdiag and the COO are the captured buffers, updated in
place between replays.  The pruned CSR (whose count reads nnz back to the
host) runs after the replay.  The library's scratch is allocated by one
eager pass before the capture, and the captured kernels clear what they need
themselves (no memset node is recorded).
"""
from __future__ import annotations

from . import device as dev


class JacobianGraph:
    def __init__(self, flow, w, coefdiag=None, bufs=None):
        t = dev.torch()
        self.flow, self.w, self.coefdiag = flow, w, coefdiag
        self.res = flow.new_state()
        self.bufs = flow.jacobian_buffers() if bufs is None else bufs
        # eager pass: allocates the library's scratch and writes the COO indices
        self._stage(write_index=True)
        t.cuda.synchronize()
        self.graph = t.cuda.CUDAGraph()
        s = t.cuda.Stream()
        s.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(s):
            with t.cuda.graph(self.graph, stream=s):
                self._stage(write_index=False)
        t.cuda.current_stream().wait_stream(s)
        t.cuda.synchronize()
        flow.ctx.sync_stream()

    def _stage(self, write_index):
        f = self.flow
        f.apply_bcs(self.w)
        f.residual(self.w, self.res)
        f.jacobian_colored(self.w, coefdiag=self.coefdiag, bufs=self.bufs, write_index=write_index)

    def replay(self):
        """BCs on w, residual into self.res, Jacobian values into self.bufs[0]."""
        self.graph.replay()
        return self.bufs
