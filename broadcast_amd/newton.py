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
    unknown u = e + 5 (j + jm i) (0-based) is permuted to outer-major blocks.

    The band half-width kb (in blocks) is the widest coupling of every block
    row but the first: on the O-mesh (j-line blocks) the wall row j = 1 also
    reaches j = 5 (the off-centred Euler flux of the face j = 2), one block
    beyond the radial stencil's 3.  Such far blocks of block row 0 are kept
    outside the band (eliminating block 0 turns them into in-band updates of
    block rows 1..kb), so the band stays 2 kb + 1 blocks wide: with kb = 3
    instead of 4, 9 instead of 16 Schur products per block step.

    With look-ahead (device tensors), block step I updates the next diagonal
    block first and factors it on a side stream while the remaining Schur
    products of step I run on the current stream."""

    def __init__(self, im, jm, outer="i", kb=None, group=1, lookahead=True, pivot=None, diag="inv"):
        self.t = dev.torch()
        # pivot=None: the project's elimination without pivoting on the device
        # for the diagonal blocks -- diag="inv": their inverses by Gauss-Jordan
        # (bca_inv_np), the block row and the solves multiplied by them;
        # diag="lu": their LU factors (bca_lu_np) and triangular solves.
        # torch's pivoted LU on the CPU or when asked (pivot=True) or when a
        # pivot comes out too small.
        self.pivot = pivot
        if diag not in ("inv", "lu"):
            raise ValueError("diag must be 'inv' or 'lu'")
        self.diag = diag
        self.im, self.jm = im, jm
        self.outer = outer
        self.n = 5 * im * jm
        lines = im if outer == "i" else jm
        line = 5 * jm if outer == "i" else 5 * im
        self.B = group * line                      # unknowns per block: `group` outer lines
        self.nb = -(-lines // group)
        self.np = self.nb * self.B                 # padded with identity rows past n
        self.kb = kb
        self.lookahead = lookahead

    def _perm(self, u):
        """Block-major index of the reference's unknown u."""
        if self.outer == "i":
            return u
        e = u % 5
        j = (u // 5) % self.jm
        i = u // (5 * self.jm)
        return e + 5 * (i + self.im * j)

    def factor(self, indptr, indices, data, shift=None):
        """Factor the CSR matrix (minus shift * I when shift is given)."""
        t = self.t
        dv = data.device
        n, B = self.n, self.B
        counts = (indptr[1:] - indptr[:-1]).long()
        rows = self._perm(t.repeat_interleave(t.arange(n, device=dv), counts))
        cols = self._perm(indices.long())
        K, J = rows // B, cols // B
        d = J - K
        if self.kb is None:
            rest = K > 0
            kb = int(d[rest].abs().max()) if bool(rest.any()) else 0
            kb = max(kb, int((-d).max()), 1 if self.nb > 1 else 0)
            self.kb = min(kb, self.nb - 1)
        kb = self.kb
        W = (2 * kb + 1) * B
        far = d > kb
        if bool((far & (K > 0)).any()) or bool((d < -kb).any()):
            raise ValueError("entry outside the block band")
        self.nfar = int(d[far].max()) - kb if bool(far.any()) else 0
        inb = ~far
        R = t.zeros((self.nb, B, W), dtype=data.dtype, device=dv)
        R.view(-1).index_put_((((K * B + rows % B) * W + (d + kb) * B + cols % B)[inb],), data[inb])
        # far blocks of block row 0: columns kb+1 .. kb+nfar
        F = t.zeros((B, self.nfar * B), dtype=data.dtype, device=dv)
        if self.nfar:
            F.view(-1).index_put_((((rows % B) * (self.nfar * B) + (d - kb - 1) * B + cols % B)[far],), data[far])
        if shift is not None:
            u = t.arange(n, device=dv)
            R.view(-1)[((u // B) * B + u % B) * W + kb * B + u % B] -= shift
        if self.np > n:
            pad = t.arange(n, self.np, device=dv)
            R.view(-1)[((pad // B) * B + pad % B) * W + kb * B + pad % B] = 1.0
        self.R, self.F = R, F
        self.lu, self.piv = [None] * self.nb, [None] * self.nb
        # block Gaussian elimination: D_I = A(I, I) after the updates of the
        # steps before; X(I, J) = D_I^-1 A(I, J) stored in place of A(I, J);
        # A(K, J) -= A(K, I) X(I, J) for K, J in I+1..I+kb
        self.np_lu = self.pivot is not True and dv.type == "cuda" and data.dtype == t.float64
        if self.np_lu:
            self._bad = t.zeros(1, dtype=t.int32, device=dv)
            if self.np_factor(R, F):
                return self
            # a pivot came out too small: refactor with partial pivoting
            self.np_lu = False
            self.pivot, self.kb = True, None
            return self.factor(indptr, indices, data, shift)
        side = t.cuda.Stream() if (self.lookahead and dv.type == "cuda") else None
        cur = t.cuda.current_stream() if side is not None else None
        ready = None
        for I in range(self.nb):
            if ready is not None:
                cur.wait_event(ready)
                ready = None
            if self.lu[I] is None:
                self.lu[I], self.piv[I], _ = t.linalg.lu_factor_ex(R[I, :, kb * B:(kb + 1) * B])
            LU, piv = self.lu[I], self.piv[I]
            hi = min(I + kb, self.nb - 1)
            if hi == I:
                continue
            Xr = R[I, :, (kb + 1) * B:(kb + 1 + hi - I) * B]
            Xr.copy_(t.linalg.lu_solve(LU, piv, Xr))
            if I == 0 and self.nfar:
                F.copy_(t.linalg.lu_solve(LU, piv, F))
            # the next diagonal block first, then its factorization beside
            # the rest of this step's updates
            R[I + 1, :, kb * B:(kb + 1) * B].sub_(R[I + 1, :, (kb - 1) * B:kb * B] @ Xr[:, :B])
            if side is not None:
                ev = t.cuda.Event()
                ev.record(cur)
                side.wait_event(ev)
                with t.cuda.stream(side):
                    A1 = R[I + 1, :, kb * B:(kb + 1) * B]
                    self.lu[I + 1], self.piv[I + 1], _ = t.linalg.lu_factor_ex(A1)
                    ready = t.cuda.Event()
                    ready.record(side)
                self.lu[I + 1].record_stream(cur)
                self.piv[I + 1].record_stream(cur)
            for dd in range(1, hi - I + 1):
                K2 = I + dd
                L = R[K2, :, (kb - dd) * B:(kb - dd + 1) * B]
                if dd == 1:
                    # the diagonal block of K2 = I+1 is done: the blocks right of it
                    if hi - I > 1:
                        R[K2, :, (kb + 1) * B:(kb + hi - I) * B].sub_(L @ Xr[:, B:])
                else:
                    R[K2, :, (kb - dd + 1) * B:(kb - dd + 1 + hi - I) * B].sub_(L @ Xr)
                if I == 0 and self.nfar:
                    # far block f (absolute block kb + f) lies at offset kb + f - K2 of row K2
                    for f in range(1, self.nfar + 1):
                        off = kb + f - K2
                        if off <= kb:
                            R[K2, :, (kb + off) * B:(kb + off + 1) * B].sub_(L @ F[:, (f - 1) * B:f * B])
        if ready is not None:
            cur.wait_event(ready)
        return self

    def np_factor(self, R, F):
        """Block elimination with the project's unpivoted kernels for the
        diagonal blocks: with diag="inv" each diagonal block is replaced by its
        inverse (bca_inv_np) and the block row X = D^-1 A(I, J) is one GEMM;
        with diag="lu" by its LU factors (bca_lu_np) and X comes from two
        triangular solves.  The first block of X and the next diagonal
        block's update go first; that block is then inverted/factored on a
        side stream while the rest of the step's products run.  False if a
        pivot was too small."""
        import ctypes as C
        from . import _lib

        t = self.t
        B, kb = self.B, self.kb
        L_ = _lib.lib()
        W = R.shape[2]
        cur = t.cuda.current_stream()
        side = t.cuda.Stream() if self.lookahead else cur
        self._bad.zero_()
        inv = self.diag == "inv"
        # scratch of each stream the diagonal blocks are processed on
        wsize = B * B if inv else 32 * 32
        work = [t.empty(wsize, dtype=t.float64, device=R.device) for _ in range(2 if self.lookahead else 1)]

        def diag_np(I, stream):
            D = R[I, :, kb * B:(kb + 1) * B]
            wk = work[1 if (self.lookahead and stream is side) else 0]
            sp = C.c_void_p(stream.cuda_stream)
            if inv:
                rc = L_.bca_inv_np(sp, C.c_void_p(D.data_ptr()), B, W, C.c_void_p(wk.data_ptr()), 1e-14,
                                   C.c_void_p(self._bad.data_ptr()))
            else:
                rc = L_.bca_lu_np(sp, C.c_void_p(D.data_ptr()), B, W, 1e-14, C.c_void_p(self._bad.data_ptr()),
                                  C.c_void_p(wk.data_ptr()))
            _lib.check(rc, "bca_inv_np" if inv else "bca_lu_np")

        def apply_inv(D, X):
            """X <- D^-1 X in place."""
            if inv:
                X.copy_(D @ X)
            else:
                X.copy_(t.linalg.solve_triangular(D, t.linalg.solve_triangular(D, X, upper=False,
                                                                               unitriangular=True), upper=True))

        diag_np(0, cur)
        ready = None
        for I in range(self.nb):
            if ready is not None:
                cur.wait_event(ready)
                ready = None
            hi = min(I + kb, self.nb - 1)
            if hi == I:
                continue
            D = R[I, :, kb * B:(kb + 1) * B]
            Xr = R[I, :, (kb + 1) * B:(kb + 1 + hi - I) * B]
            apply_inv(D, Xr[:, :B])
            R[I + 1, :, kb * B:(kb + 1) * B].sub_(R[I + 1, :, (kb - 1) * B:kb * B] @ Xr[:, :B])
            if self.lookahead:
                ev = t.cuda.Event()
                ev.record(cur)
                side.wait_event(ev)
                diag_np(I + 1, side)
                ready = t.cuda.Event()
                ready.record(side)
            else:
                diag_np(I + 1, cur)
            if hi - I > 1:
                apply_inv(D, Xr[:, B:])
            if I == 0 and self.nfar:
                apply_inv(D, F)
            for dd in range(1, hi - I + 1):
                K2 = I + dd
                L = R[K2, :, (kb - dd) * B:(kb - dd + 1) * B]
                if dd == 1:
                    if hi - I > 1:
                        R[K2, :, (kb + 1) * B:(kb + hi - I) * B].sub_(L @ Xr[:, B:])
                else:
                    R[K2, :, (kb - dd + 1) * B:(kb - dd + 1 + hi - I) * B].sub_(L @ Xr)
                if I == 0 and self.nfar:
                    for f in range(1, self.nfar + 1):
                        off = kb + f - K2
                        if off <= kb:
                            R[K2, :, (kb + off) * B:(kb + off + 1) * B].sub_(L @ F[:, (f - 1) * B:f * B])
        if ready is not None:
            cur.wait_event(ready)
        return int(self._bad.item()) == 0

    def _solve_diag(self, I, v):
        """D_I^-1 v for one block (v: (B,) or (B, k))."""
        t = self.t
        if not self.np_lu:
            return t.linalg.lu_solve(self.lu[I], self.piv[I], v.unsqueeze(1) if v.dim() == 1 else v).reshape(v.shape)
        kb, B = self.kb, self.B
        D = self.R[I, :, kb * B:(kb + 1) * B]
        if self.diag == "inv":
            return D @ v
        m = v.unsqueeze(1) if v.dim() == 1 else v
        y = t.linalg.solve_triangular(D, t.linalg.solve_triangular(D, m, upper=False, unitriangular=True), upper=True)
        return y.reshape(v.shape)

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
            y[I] = self._solve_diag(I, y[I])
            for dd in range(1, min(kb, nb - 1 - I) + 1):
                y[I + dd] -= R[I + dd, :, (kb - dd) * B:(kb - dd + 1) * B] @ y[I]
        for I in range(nb - 1, -1, -1):
            for dd in range(1, min(kb, nb - 1 - I) + 1):
                y[I] -= R[I, :, (kb + dd) * B:(kb + dd + 1) * B] @ y[I + dd]
            if I == 0 and self.nfar:
                for f in range(1, self.nfar + 1):
                    if kb + f < nb:
                        y[0] -= self.F[:, (f - 1) * B:f * B] @ y[kb + f]
        y = y.reshape(-1)[:self.n]
        if self.outer == "i":
            return y
        return y.reshape(self.jm, self.im, 5).permute(1, 0, 2).reshape(-1)


def csr_residual(csr, x, b):
    """b - A x for the CSR (indptr, indices, data) on the device."""
    import ctypes as C
    from . import _lib

    t = dev.torch()
    indptr, indices, data = csr
    y = t.empty_like(b)
    _lib.check(_lib.lib().bca_csr_spmv(dev.current_stream_ptr(), b.numel(), C.c_void_p(indptr.data_ptr()),
                                       C.c_void_p(indices.data_ptr()), C.c_void_p(data.data_ptr()),
                                       C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())), "bca_csr_spmv")
    return b - y


def refine(lu, csr, b, x, rtol=1e-12, steps=3):
    """Iterative refinement of x (in place) for A x = b with the factors of
    lu, until ||b - A x|| <= rtol ||b||, a correction no longer halves the
    residual (its rounding floor), or `steps` corrections: the guard of the
    unpivoted factorization.  Returns (corrections, final relative
    residual)."""
    bn = float(b.norm())
    k = 0
    r = csr_residual(csr, x, b)
    rel = float(r.norm()) / max(bn, 1e-300)
    while rel > rtol and k < steps:
        dx = lu.solve(r)
        x += dx
        k += 1
        r = csr_residual(csr, x, b)
        new = float(r.norm()) / max(bn, 1e-300)
        if new > 0.5 * rel:
            rel = new
            break
        rel = new
    return k, rel


@dataclass
class NewtonLog:
    norms: list = field(default_factory=list)      # max_e ||R_e||_2 before each update
    relax: list = field(default_factory=list)      # r of each iteration
    seconds: list = field(default_factory=list)    # wall time of each iteration
    split: list = field(default_factory=list)      # per update: (jacobian, factor, solve) seconds
    refinements: list = field(default_factory=list)  # per update: (refinement steps, final relative residual)
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
    outer = "j" if case.kind == "cyl" else "i"
    norm0 = ninf0 = None
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
        # the colored Jacobian straight into the pruned CSR; from the second
        # iteration on only the values are rewritten while every row keeps
        # its kept set (the same matrix as jacobian_colored + csr)
        indptr, indices, data = flow.jacobian_csr(w, coefdiag=coefdiag, refresh=True)
        t.cuda.synchronize()
        t1 = time.perf_counter()
        lu = BlockBandLU(flow.im, flow.jm, outer=outer, group=group).factor(indptr, indices, data)
        t.cuda.synchronize()
        t2 = time.perf_counter()
        rhs = interior(flow, res).reshape(-1)
        dw = lu.solve(rhs)
        log.refinements.append(refine(lu, (indptr, indices, data), rhs, dw))
        del lu
        w_prev, l2_prev = w.clone(), l2
        wi = w.view(5, flow.nj, flow.ni)[:, 3:-3, 3:-3]
        wi += dw.reshape(flow.im, flow.jm, 5).permute(2, 1, 0)
        t.cuda.synchronize()
        t3 = time.perf_counter()
        log.split.append((t1 - t0, t2 - t1, t3 - t2))
        log.seconds.append(t3 - t0)
    log.cfl = cfl
    return log


def continuation(flow_at, target, start, w=None, steps=4, ite=40, cfl=1e21, tol=1e-10, min_step=None, log=None,
                 verbose=False):
    """Continuation Newton in one flow parameter (the Reynolds number here):
    Newton at `start` (from w, or the free stream of flow_at(start)), then at
    parameters stepping to `target`, each solve starting from the previous
    baseflow; a step that does not converge is halved (down to min_step).
    flow_at(p) gives the Flow at parameter p on the same mesh.  Returns
    (flow at target, w at target, list of (p, NewtonLog)).  No reference
    counterpart: the reference's drivers start every solve from their own
    initial state."""
    from .flow import to_device

    t = dev.torch()
    f = flow_at(start)
    w = to_device(f.case.w) if w is None else w
    hist = []
    lg = newton(f, w, ite=ite, cfl=cfl, tol=tol)
    hist.append((start, lg))
    if not lg.converged:
        raise RuntimeError(f"continuation: no converged start at {start}")
    p = start
    dp = (target - start) / max(steps, 1)
    min_step = abs(dp) / 16 if min_step is None else min_step
    while (target - p) * np.sign(dp) > 1e-12:
        q = p + dp
        if (target - q) * np.sign(dp) < 1e-9 * abs(dp):
            q = target
        fq = flow_at(q)
        wq = w.clone()
        lq = newton(fq, wq, ite=ite, cfl=cfl, tol=tol)
        hist.append((q, lq))
        if verbose:
            print(f"continuation {p:.3g} -> {q:.3g}: {len(lq.norms) - 1} Newton steps, converged {lq.converged}",
                  flush=True)
        if lq.converged:
            p, w, f = q, wq, fq
        else:
            dp /= 2.0
            if abs(dp) < min_step:
                raise RuntimeError(f"continuation stalled at {p} (step {dp})")
        del wq
    t.cuda.synchronize()
    return f, w, hist
