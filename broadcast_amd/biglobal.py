"""Global stability of a spanwise-periodic baseflow on the device: the
eigenvalues of J(beta) nearest a target by shift-invert Arnoldi.

The reference (biglobal.py:59-100) composes Jac3D = Jacvol - beta^2 Dzz +
i beta Dz from the NPZ tree's CSR arrays, wraps (Jac3D - target I)^-1 as a
shell matrix (createShiftInvert, misc/PETSc_func.py:1516-1535: MUMPS LU of
Jac3D - target I) and asks SLEPc for its largest-magnitude eigenvalues
(eigPetsc3, :1569), eigenvalue = target + 1 / theta.  Here:

* J(beta) comes from the device path (Flow.jbeta, complex128 CSR);
* Jac3D - target I is factored by the block-banded LU of the Newton driver
  (broadcast_amd.newton.BlockBandLU, i-line blocks, complex with partial
  pivoting);
* the eigen solve is a restarted Arnoldi on the shift-invert operator with
  Ritz-vector restarts (SLEPc's Krylov-Schur with nev = 1 does the same
  projection with implicit restarts); the eigenvalue's true residual
  ||J x - lambda x|| / (||J||_1 ||x||) is reported, like the reference's
  setTrueResidual(True).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import device as dev
from .newton import BlockBandLU


def csr_matvec(indptr, indices, values, x):
    """y = A x for a CSR matrix on the device (any dtype)."""
    t = dev.torch()
    n = indptr.numel() - 1
    A = t.sparse_csr_tensor(indptr.long(), indices.long(), values, size=(n, n))
    return (A @ x.unsqueeze(1)).squeeze(1)


class ShiftInvert:
    """(J - target I)^-1 for a CSR J of a flow's mesh (the reference's
    createShiftInvert), by the block-banded LU."""

    def __init__(self, flow, indptr, indices, values, target):
        outer = "j" if flow.case.kind == "cyl" else "i"
        self.lu = BlockBandLU(flow.im, flow.jm, outer=outer, pivot=True).factor(indptr, indices, values,
                                                                               shift=complex(target))

    def solve(self, b):
        return self.lu.solve(b)


@dataclass
class Eigen:
    value: complex
    vector: object          # (n,) complex128 device tensor, unit 2-norm, the reference's unknown order
    residual: float         # ||J x - lambda x|| / (||J||_1 ||x||)
    restarts: int
    ritz_residual: float


def arnoldi_shift_invert(op, n, target, m=30, tol=1e-10, maxit=30, seed=0, dtype=None, device="cuda"):
    """Eigenvalue of largest |theta| of op (= (J - target I)^-1): restarted
    Arnoldi with m Krylov vectors, restarted from the Ritz vector until the
    Ritz residual |h_{m+1,m} y_m| <= tol |theta|.  Returns (lambda =
    target + 1/theta, Ritz vector, restarts, Ritz residual)."""
    t = dev.torch()
    dtype = t.complex128 if dtype is None else dtype
    g = t.Generator(device=device).manual_seed(seed)
    v = t.complex(t.randn(n, dtype=t.float64, device=device, generator=g),
                  t.randn(n, dtype=t.float64, device=device, generator=g)).to(dtype)
    v /= v.norm()
    for rs in range(maxit):
        V = t.zeros((m + 1, n), dtype=dtype, device=device)
        H = np.zeros((m + 1, m), dtype=np.complex128)
        V[0] = v
        k_used = m
        for k in range(m):
            wv = op(V[k])
            # classical Gram-Schmidt, twice
            for _ in range(2):
                h = (V[:k + 1].conj() @ wv)
                wv = wv - h @ V[:k + 1]
                H[:k + 1, k] += h.cpu().numpy()
            H[k + 1, k] = float(wv.norm())
            if H[k + 1, k] < 1e-300:
                k_used = k + 1
                break
            V[k + 1] = wv / H[k + 1, k]
        Hm = H[:k_used, :k_used]
        th, Y = np.linalg.eig(Hm)
        i = int(np.argmax(np.abs(th)))
        y = Y[:, i] / np.linalg.norm(Y[:, i])
        ritz_res = abs(H[k_used, k_used - 1] * y[-1]) if k_used < m + 1 and k_used == m else 0.0
        x = t.as_tensor(y, dtype=dtype, device=device) @ V[:k_used]
        x /= x.norm()
        if ritz_res <= tol * abs(th[i]):
            return target + 1.0 / th[i], x, rs, ritz_res / abs(th[i])
        v = x
    return target + 1.0 / th[i], x, maxit, ritz_res / abs(th[i])


def eig_near(flow, indptr, indices, values, target, m=30, tol=1e-10, maxit=30, seed=0):
    """The eigenvalue of the CSR J (J(beta) of this flow's mesh; flow gives
    im, jm and case.kind) nearest target in the shift-invert sense, its
    eigenvector and true residual.  Device or CPU tensors."""
    t = dev.torch()
    n = indptr.numel() - 1
    si = ShiftInvert(flow, indptr, indices, values, target)
    lam, x, rs, rr = arnoldi_shift_invert(si.solve, n, target, m=m, tol=tol, maxit=maxit, seed=seed,
                                          dtype=values.dtype, device=values.device)
    Jx = csr_matvec(indptr, indices, values, x)
    # ||J||_1: largest column sum of |J|
    colsum = t.zeros(n, dtype=t.float64, device=values.device).index_add_(0, indices.long(), values.abs())
    res = float((Jx - lam * x).norm()) / (float(colsum.max()) * float(x.norm()))
    return Eigen(complex(lam), x, res, rs, float(rr))
