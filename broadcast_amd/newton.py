"""Newton baseflow driver on the device (BROADCAST_npz.py:985-1196).

Per iteration, as the reference's `fixed_point` mode:

* boundary conditions, residual and its norms (compute_norml2inf);
* relaxation r = max over the first three variables of the norm ratios to
  the first iteration (:1062), coefdiag = r / dt * vol with
  dt = cfl (yc(1,2) - yc(1,1)) / (1/M + 1) (:1011-1012, :1063-1066), on the
  O-mesh dt = cfl |xc(1,2) - xc(1,1)| / (1/M + 1) (cylinder.py:842-894);
* the colored Jacobian -dR/dw + coefdiag, pruned, in CSR (:1068-1150);
* the linear solve J dw = R (MUMPS LU in the reference, :1152-1157; here the
  device block-banded LU below);
* w += dw.

Departures, each on the side of robustness: the reference runs a fixed
number of iterations, `newton` also stops when the residual norm
max_e ||R_e||_2 (compute_norml2, srcfv/norm.F90:2-32) falls below `tol`; on
a NaN step the reference subtracts the previous step and halves a cfl it no
longer uses (:1167-1172), `newton` restores the state before the step and
divides the cfl used for dt by `cfl_cut`, also when the residual norm grows
by more than `growth`.

Block-banded LU.  In the reference's unknown order (C-order ravel of
(im, jm, 5), i outermost) the Jacobian couples i-lines at most 3 apart, so it
is block banded with blocks of one i-line (5 jm unknowns).  The factorization
is block Gaussian elimination: each diagonal block by LU with partial
pivoting (rocSOLVER getrf through torch.linalg.lu_factor), the row blocks to
its right by one solve with it, the band below updated by one GEMM per block
row (rocBLAS); no fill leaves the band.  Storage is dense per block row,
R[K] = [A(K, K-kb) ... A(K, K+kb)], (nblocks, B, (2 kb + 1) B) doubles.
An i-periodic mesh (the O-mesh) couples the first and last i-lines; it is
solved with j outermost instead (blocks of one j-line), where the band is
that of the radial stencil.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import device as dev


class BlockBandLU:
    """Dense block-banded LU on the device of an n x n CSR matrix whose
    unknown u = e + 5 (j + jm i) (0-based) is permuted to outer-major blocks."""

    def __init__(self, im, jm, outer="i", kb=None, group=1):
        self.t = dev.torch()
        self.im, self.jm = im, jm
        self.outer = outer
        self.n = 5 * im * jm
        lines = im if outer == "i" else jm
        line = 5 * jm if outer == "i" else 5 * im
        self.B = group * line                      # unknowns per block: `group` outer lines
        self.nb = -(-lines // group)
        self.np = self.nb * self.B                 # padded with identity rows past n
        self.kb = kb

    def _perm(self, u):
        """Block-major index of the reference's unknown u."""
        if self.outer == "i":
            return u
        e = u % 5
        j = (u // 5) % self.jm
        i = u // (5 * self.jm)
        return e + 5 * (i + self.im * j)

    def factor(self, indptr, indices, data):
        t = self.t
        dv = data.device
        n, B = self.n, self.B
        counts = (indptr[1:] - indptr[:-1]).long()
        rows = self._perm(t.repeat_interleave(t.arange(n, device=dv), counts))
        cols = self._perm(indices.long())
        K, J = rows // B, cols // B
        kb = int((J - K).abs().max()) if self.kb is None else self.kb
        self.kb = kb
        W = (2 * kb + 1) * B
        d = J - K + kb
        if bool(((d < 0) | (d > 2 * kb)).any()):
            raise ValueError("entry outside the block band")
        R = t.zeros((self.nb, B, W), dtype=data.dtype, device=dv)
        R.view(-1).index_put_(((K * B + rows % B) * W + d * B + cols % B,), data)
        if self.np > n:
            pad = t.arange(n, self.np, device=dv)
            R.view(-1)[((pad // B) * B + pad % B) * W + kb * B + pad % B] = 1.0
        self.R = R
        self.lu, self.piv = [], []
        # block Gaussian elimination: D_I = A(I, I) after the updates of the
        # steps before; X(I, J) = D_I^-1 A(I, J) stored in place of A(I, J);
        # A(K, J) -= A(K, I) X(I, J) for K, J in I+1..I+kb
        for I in range(self.nb):
            LU, piv, _ = t.linalg.lu_factor_ex(R[I, :, kb * B:(kb + 1) * B])
            self.lu.append(LU)
            self.piv.append(piv)
            hi = min(I + kb, self.nb - 1)
            if hi == I:
                continue
            Xr = R[I, :, (kb + 1) * B:(kb + 1 + hi - I) * B]
            Xr.copy_(t.linalg.lu_solve(LU, piv, Xr))
            for dd in range(1, hi - I + 1):
                K2 = I + dd
                R[K2, :, (kb - dd + 1) * B:(kb - dd + 1 + hi - I) * B].sub_(R[K2, :, (kb - dd) * B:(kb - dd + 1) * B] @ Xr)
        return self

    def solve(self, b):
        """x with A x = b, b in the reference's unknown order."""
        t = self.t
        B, kb, nb = self.B, self.kb, self.nb
        if self.outer == "i":
            y = b.reshape(-1)
        else:
            y = b.reshape(self.im, self.jm, 5).permute(1, 0, 2).reshape(-1)
        y = t.cat((y, y.new_zeros(self.np - self.n))).reshape(nb, B)
        R = self.R
        for I in range(nb):
            y[I] = t.linalg.lu_solve(self.lu[I], self.piv[I], y[I].unsqueeze(1)).squeeze(1)
            for dd in range(1, min(kb, nb - 1 - I) + 1):
                y[I + dd] -= R[I + dd, :, (kb - dd) * B:(kb - dd + 1) * B] @ y[I]
        for I in range(nb - 1, -1, -1):
            for dd in range(1, min(kb, nb - 1 - I) + 1):
                y[I] -= R[I, :, (kb + dd) * B:(kb + dd + 1) * B] @ y[I + dd]
        y = y.reshape(-1)[:self.n]
        if self.outer == "i":
            return y
        return y.reshape(self.jm, self.im, 5).permute(1, 0, 2).reshape(-1)


@dataclass
class NewtonLog:
    norms: list = field(default_factory=list)      # max_e ||R_e||_2 before each update
    relax: list = field(default_factory=list)      # r of each iteration
    seconds: list = field(default_factory=list)    # wall time of each iteration
    rollbacks: int = 0
    converged: bool = False


def interior(flow, a):
    """(im, jm, 5) view-copy of the interior of a flat state, the reference's
    unknown order after reshape(-1)."""
    return a.view(5, flow.nj, flow.ni)[:, 3:-3, 3:-3].permute(2, 1, 0).contiguous()


def newton(flow, w, ite=10, cfl=1e10, tol=1e-10, log=None, verbose=False, growth=1e3, cfl_cut=10.0, group=1):
    """The reference's Newton iteration on the flat device state w (updated
    in place).  Returns a NewtonLog."""
    t = dev.torch()
    log = NewtonLog() if log is None else log
    case = flow.case
    mach = case.phys.mach
    if case.kind == "cyl":
        # cylinder.py:842: |xc(1,2) - xc(1,1)|, the radial spacing at i = 1
        xc = flow.xc.view(flow.nj, flow.ni)
        dy = abs(float(xc[4, 3] - xc[3, 3]))
    else:
        # BROADCAST_npz.py:1011: yc(1,2) - yc(1,1), the first cell height
        yc = flow.yc.view(flow.nj, flow.ni)
        dy = float(yc[4, 3] - yc[3, 3])
    vol_int = flow.vol.view(flow.nj, flow.ni)[3:-3, 3:-3]      # (jm, im)
    res = flow.new_state()
    bufs = flow.jacobian_buffers()
    outer = "j" if case.kind == "cyl" else "i"
    norm0 = ninf0 = None
    first = True
    w_prev, l2_prev = None, None
    it = 0
    while it < ite:
        it += 1
        t0 = time.perf_counter()
        flow.apply_bcs(w)
        res.zero_()
        flow.residual(w, res)
        norm, ninf = flow.norms(res)   # per variable sqrt(sum R^2) and (sum R^10)^(1/10)
        l2 = float(np.max(norm))
        if w_prev is not None and (not np.isfinite(l2) or l2 > growth * l2_prev):
            # rejected step: back to the previous state with a smaller cfl
            w.copy_(w_prev)
            cfl /= cfl_cut
            log.rollbacks += 1
            if verbose:
                print(f"newton it {it}: ||R|| = {l2:.3e} rejected, cfl -> {cfl:.3e}", flush=True)
            w_prev = None
            continue
        log.norms.append(l2)
        if verbose:
            print(f"newton it {it}: max ||R||_2 = {l2:.3e}", flush=True)
        if tol is not None and l2 < tol:
            log.converged = True
            log.seconds.append(time.perf_counter() - t0)
            break
        if norm0 is None:
            norm0 = 1.0 / np.maximum(norm, 1e-15)
            ninf0 = 1.0 / np.maximum(ninf, 1e-15)
        r = float(np.max([norm[:3] * norm0[:3], ninf[:3] * ninf0[:3]]))
        log.relax.append(r)
        dtm1 = (1.0 / mach + 1.0) / (cfl * dy)
        coefdiag = (r * dtm1 * vol_int).contiguous().reshape(-1)   # (im, jm) in Fortran order
        jac, ia, ja = flow.jacobian_colored(w, coefdiag=coefdiag, bufs=bufs, write_index=first)
        first = False
        indptr, indices, data = flow.csr(jac)
        lu = BlockBandLU(flow.im, flow.jm, outer=outer, group=group).factor(indptr, indices, data)
        dw = lu.solve(interior(flow, res).reshape(-1))
        w_prev, l2_prev = w.clone(), l2
        wi = w.view(5, flow.nj, flow.ni)[:, 3:-3, 3:-3]
        wi += dw.reshape(flow.im, flow.jm, 5).permute(2, 1, 0)
        t.cuda.synchronize()
        log.seconds.append(time.perf_counter() - t0)
    log.cfl = cfl
    return log
