"""Eigenvalues of J(beta) near a target on the device (broadcast_amd.biglobal:
J(beta) from the GPU path, shift-invert by the block-banded LU, restarted
Arnoldi) against scipy's shift-invert eigs on the same CSR copied to the
host -- the reference's biglobal.py:59-100 pipeline."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from broadcast_amd.cases import bl_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


@pytest.mark.parametrize("beta,target", [(0.5, -0.1), (0.25, 0.05 + 0.2j)])
def test_eig_near_target_matches_scipy(torch_mod, beta, target):
    from broadcast_amd.biglobal import eig_near
    from broadcast_amd.flow import Flow, to_device

    case = bl_case(40, 24, seed=91)
    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    jac, _, _ = f.jacobian_colored(w)
    jdz, jdz2, _, _ = f.dz_colored(w)
    indptr, indices, vals = f.jbeta(jac, jdz, jdz2, [beta])
    v = vals[0].contiguous()
    n = indptr.numel() - 1
    A = sp.csr_matrix((v.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
    ref = spla.eigs(A.tocsc(), k=1, sigma=target, which="LM", return_eigenvectors=False)[0]
    e = eig_near(f, indptr, indices, v, target, m=30, tol=1e-12)
    print(f"beta {beta} target {target}: lambda {e.value:.10g} (scipy {ref:.10g}), residual {e.residual:.2e}, "
          f"restarts {e.restarts}", flush=True)
    assert abs(e.value - ref) <= 1e-8 * max(abs(ref), 1.0)
    assert e.residual < 1e-12
