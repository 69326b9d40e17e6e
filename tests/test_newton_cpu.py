"""The block-banded LU of the Newton driver (broadcast_amd.newton), run with
torch on the CPU against scipy's sparse direct solve, on the oracle's colored
Jacobians: the boundary layer with i-line blocks, the O-mesh (i-periodic)
with j-line blocks."""
import numpy as np
import pytest

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case, cylinder_case


@pytest.fixture(scope="module")
def O():
    return oracle.Oracle()


@pytest.mark.parametrize("mk,outer,group", [(lambda: bl_case(16, 10, seed=21), "i", 1),
                                            (lambda: bl_case(16, 10, seed=21), "i", 3),
                                            (lambda: bl_case(12, 9, seed=22, reynolds=3e5), "j", 1),
                                            (lambda: cylinder_case(14, 9, seed=23), "j", 2)])
def test_block_band_lu_matches_sparse_solve(O, mk, outer, group):
    import scipy.sparse.linalg as spla
    import torch
    from broadcast_amd.newton import BlockBandLU

    case = mk()
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    rng = np.random.default_rng(1)
    cd = np.asfortranarray(rng.uniform(0.0, 1e-3, (case.im, case.jm)))
    J = D.colored_jacobian(O, case, g, w, coefdiag=cd)
    J.sort_indices()
    b = rng.standard_normal(J.shape[0])
    ref = spla.spsolve(J.tocsc(), b)
    lu = BlockBandLU(case.im, case.jm, outer=outer, group=group).factor(torch.from_numpy(J.indptr.astype(np.int32)),
                                                           torch.from_numpy(J.indices.astype(np.int32)),
                                                           torch.from_numpy(J.data.copy()))
    x = lu.solve(torch.from_numpy(b)).numpy()
    # normwise backward error, next to that of the sparse direct solve (the
    # wall rows make the Jacobian badly scaled: entries span ~10 decades)
    anorm = abs(J).sum(axis=1).max()
    berr = np.linalg.norm(J @ x - b, np.inf) / (anorm * np.linalg.norm(x, np.inf) + np.linalg.norm(b, np.inf))
    berr_ref = np.linalg.norm(J @ ref - b, np.inf) / (anorm * np.linalg.norm(ref, np.inf) + np.linalg.norm(b, np.inf))
    assert berr <= max(10 * berr_ref, 1e-15), (berr, berr_ref)
    assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref)
    assert lu.kb <= -(-3 // group)
