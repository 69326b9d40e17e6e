"""Shift-invert Arnoldi of broadcast_amd.biglobal on the CPU (torch CPU
tensors) against scipy's shift-invert eigs, on J(beta) = Jac/vol - beta^2 Dzz
+ i beta Dz composed from the oracle's colored Jacobian and Dz/Dzz (the
reference's biglobal.py:82 on the NPZ arrays)."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case


def jbeta_oracle(O, case, beta):
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    J = D.colored_jacobian(O, case, g, w)
    jdz, jdz2, ia, ja = D.dz_colored(O, case, g, w)
    n = 5 * case.im * case.jm
    keep = (np.abs(jdz) > 2e-16) | (np.abs(jdz2) > 2e-16)
    Dz = sp.csr_matrix((jdz[keep], (ia[keep], ja[keep])), shape=(n, n))
    Dzz = sp.csr_matrix((jdz2[keep], (ia[keep], ja[keep])), shape=(n, n))
    vol = np.ravel(g["vol"][case.gh:-case.gh, case.gh:-case.gh], order="C")  # (im, jm) C-order = cell order
    Jvol = sp.diags(np.repeat(1.0 / vol, 5)) @ J
    return (Jvol - beta ** 2 * Dzz + 1j * beta * Dz).tocsr()


@pytest.mark.parametrize("beta,target", [(0.5, -0.1), (0.0, 0.2 + 0.1j)])
def test_shift_invert_arnoldi_matches_scipy(beta, target):
    import torch
    from broadcast_amd.biglobal import eig_near

    O = oracle.Oracle()
    case = bl_case(12, 9, seed=3)
    A = jbeta_oracle(O, case, beta)
    A.sort_indices()
    ref = spla.eigs(A.tocsc(), k=1, sigma=target, which="LM", return_eigenvectors=False)[0]
    flow = SimpleNamespace(im=case.im, jm=case.jm, case=case)
    e = eig_near(flow, torch.from_numpy(A.indptr.astype(np.int32)), torch.from_numpy(A.indices.astype(np.int32)),
                 torch.from_numpy(A.data.astype(np.complex128)), target, m=20, tol=1e-12)
    assert abs(e.value - ref) <= 1e-8 * max(abs(ref), 1.0), (e.value, ref)
    assert e.residual < 1e-12
