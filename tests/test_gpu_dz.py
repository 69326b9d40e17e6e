"""Spanwise operators Dz, Dzz (bca_dz_5p, bca_dz_colored) and the J(beta)
batch (bca_jbeta_*) against the reference's vectors and the oracle.

FP64; operators to 1e-12 normwise per variable (the project's parity bound), COO
indices bit-exact, J(beta) from identical COO inputs exact up to the sign of
zeros, and at the 512x192 boundary layer with 16 wavenumbers through the
structure J(beta) = J(0) - beta^2 Dzz + i beta Dz.
"""
import numpy as np
import pytest

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case
from golden_util import load, case_from, close

pytestmark = pytest.mark.gpu

CASES = ["bl24x16", "bl13x9", "cyl28x12"]
GEO = ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


@pytest.fixture(scope="module")
def O():
    return oracle.Oracle()


def _dev(a):
    from broadcast_amd.flow import to_device

    return to_device(a)


@pytest.mark.parametrize("name", CASES)
def test_dz_dropin_matches_reference(torch_mod, name):
    from broadcast_amd.srcfv import f_dz

    fx = load(name)
    case = case_from(fx)
    g = {k: np.asfortranarray(fx[f"geo_{k}"]) for k in GEO}
    a = case.phys.scheme_args()[:9]
    for fn, key in ((f_dz.coeffs_5p_dz, "dz"), (f_dz.coeffs_5p_dz2, "dz2")):
        out = np.zeros_like(fx["w_bc"], order="F")
        fn(out, np.asfortranarray(fx["w_bc"]), np.asfortranarray(fx["wd_bc"][0]), g["x0"], g["y0"], g["nx"],
           g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], case.gh, *a, case.im, case.jm)
        ok, err = close(out, fx[key], 1e-12)
        assert ok, (name, key, err)


def test_dz_closed_forms(torch_mod, O):
    """As the oracle's test: uniform rho = 1, u = U state; Dz e_(rho w) =
    (1, u, 0, 0, H), Dzz e_(rho u) = (0, -mu, 0, 0, mu U (gamma/Pr - 1))."""
    from broadcast_amd.srcfv import f_dz

    case = bl_case(12, 9, seed=0, ratio=1.1, dy1=1e-2)
    g = D.geometry(O, case)
    U = 0.3
    w = np.zeros((case.ni, case.nj, 5), order="F")
    w[:, :, 0] = 1.0
    w[:, :, 1] = U
    w[:, :, 4] = case.phys.p_inf / (case.phys.gam - 1.0) + 0.5 * U * U
    a = case.phys.scheme_args()[:9]
    e4 = np.zeros_like(w, order="F")
    e4[:, :, 3] = 1.0
    dz = np.zeros_like(w, order="F")
    f_dz.coeffs_5p_dz(dz, w, e4, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], 3, *a,
                      case.im, case.jm)
    H = w[0, 0, 4] + case.phys.p_inf
    np.testing.assert_allclose(dz[3:-3, 3:-3].reshape(-1, 5), np.tile([1.0, U, 0.0, 0.0, H], (case.im * case.jm, 1)),
                               rtol=1e-14, atol=1e-14)
    e2 = np.zeros_like(w, order="F")
    e2[:, :, 1] = 1.0
    dz2 = np.zeros_like(w, order="F")
    f_dz.coeffs_5p_dz2(dz2, w, e2, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], 3, *a,
                       case.im, case.jm)
    ref = np.zeros_like(dz2)
    O.coeffs_5p_dz2(ref, w, e2, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], 3, *a,
                    case.im, case.jm)
    np.testing.assert_allclose(dz2, ref, rtol=1e-13, atol=1e-16)


def test_dz_scatter_matches_oracle(torch_mod, O):
    from broadcast_amd.misc import f_misc

    im, jm, gh = 15, 11, 3
    nb = im * jm * 49 * 25
    rng = np.random.default_rng(2)
    dz = np.asfortranarray(rng.standard_normal((im + 6, jm + 6, 5)))
    g = [np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)]
    o = [np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)]
    for m, l, k in [(0, 0, 0), (2, 5, 3), (4, 6, 6), (1, 3, 0), (3, 2, 5), (0, 4, 4)]:
        f_misc.computejacobianfromdz(*g, dz, m, l, k, gh, im, jm)
        O.computejacobianfromdz(*o, dz, m, l, k, gh, im, jm)
    for x, y in zip(g, o):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("mk", [lambda: bl_case(29, 18, seed=51), lambda: bl_case(20, 13, seed=52, reynolds=3e5)])
def test_dz_colored_matches_oracle_loop(torch_mod, O, mk):
    """The whole Dz and Dzz COOs of the batched call against the oracle's seed
    loop (testvector + tangent BCs + coeffs_5p_dz[2] + computejacobianfromdz),
    then J(beta) on the oracle's COOs against the same composition in numpy."""
    from broadcast_amd.flow import Flow
    import scipy.sparse as sp

    case = mk()
    print(f"case bl {case.im}x{case.jm}", flush=True)
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    jdz_o, jdz2_o, ia_o, ja_o = D.dz_colored(O, case, g, w)
    f = Flow(case)
    wg = _dev(case.w)
    f.apply_bcs(wg)
    jdz, jdz2, ia, ja = (x.cpu().numpy() for x in f.dz_colored(wg))
    np.testing.assert_array_equal(ia, ia_o)
    np.testing.assert_array_equal(ja, ja_o)
    n = 5 * case.im * case.jm
    for got, ref in ((jdz, jdz_o), (jdz2, jdz2_o)):
        rowmax = np.zeros(n)
        np.maximum.at(rowmax, ia_o, np.abs(ref))
        err = np.abs(got - ref)
        bad = err > 1e-12 * np.maximum(rowmax[ia_o], 1e-300)
        assert not bad.any(), (int(bad.sum()), float((err / np.maximum(rowmax[ia_o], 1e-300)).max()))
    # J(beta) from the oracle's COOs
    rng = np.random.default_rng(3)
    cd = np.asfortranarray(rng.uniform(0.5, 1.5, (case.im, case.jm)))
    jac_o, jia, jja = D.colored_jacobian(O, case, g, w, coefdiag=cd, coo=True)
    betas = [0.0, 1.0 / 15.0, 0.4, 1.0]
    indptr, indices, vals = f.jbeta(_dev(jac_o), _dev(jdz_o), _dev(jdz2_o), betas)
    indptr, indices, vals = indptr.cpu().numpy(), indices.cpu().numpy(), vals.cpu().numpy()
    mini = 2e-16
    vol = g["vol"]
    kx, ky, ky2 = np.abs(jac_o) > mini, np.abs(jdz_o) > mini, np.abs(jdz2_o) > mini
    rows_cell_i, rows_cell_j = jia // (case.jm * 5), (jia % (case.jm * 5)) // 5
    jv = np.where(kx, jac_o / vol[rows_cell_i + 3, rows_cell_j + 3], 0.0)
    keep = kx | ky | ky2
    order = np.lexsort((jja[keep], jia[keep]))
    r_rows, r_cols = jia[keep][order], jja[keep][order]
    np.testing.assert_array_equal(indptr, np.concatenate(([0], np.cumsum(np.bincount(r_rows, minlength=n)))))
    np.testing.assert_array_equal(indices, r_cols)
    for q, b in enumerate(betas):
        re = (jv - (b * b) * np.where(ky2, jdz2_o, 0.0))[keep][order]
        im_ = (b * np.where(ky, jdz_o, 0.0))[keep][order]
        np.testing.assert_array_equal(vals[q].real, re)
        np.testing.assert_array_equal(vals[q].imag, im_)
    # and the scipy/PETSc-style sum of the three pruned matrices, densely
    J = sp.csr_matrix((jv[kx], (jia[kx], jja[kx])), shape=(n, n))
    Dz = sp.csr_matrix((jdz_o[ky], (ia_o[ky], ja_o[ky])), shape=(n, n))
    Dzz = sp.csr_matrix((jdz2_o[ky2], (ia_o[ky2], ja_o[ky2])), shape=(n, n))
    b = betas[2]
    ref = (J - b ** 2 * Dzz + b * 1j * Dz).toarray()
    got = sp.csr_matrix((vals[2], indices, indptr), shape=(n, n)).toarray()
    assert np.array_equal(got, ref)


def test_jbeta_full_size(torch_mod):
    """512x192 boundary layer (M 0.5, Re_L 1e5), 16 wavenumbers beta_k = k/15:
    Jacobian, Dz and Dzz assembled on the device.  Exact checks: on every
    entry, Im J(beta_k) = beta_k Im J(1) (beta_15 = 1 gives Dz itself) and
    J(0) is real; on 300 sampled rows, every J(beta_k) row equals the union of
    the rows of the three separately pruned CSRs (the device prune + CSR of
    each COO) combined as Jac/vol - beta^2 Dzz + i beta Dz in numpy."""
    torch = torch_mod
    from broadcast_amd.flow import Flow

    case = bl_case(512, 192, seed=2000)
    print(f"case bl {case.im}x{case.jm} J(beta) x16", flush=True)
    f = Flow(case)
    w = _dev(case.w)
    f.apply_bcs(w)
    cd = torch.full((case.im * case.jm,), 1e-2, dtype=torch.float64, device="cuda")
    jac, _, _ = f.jacobian_colored(w, coefdiag=cd)
    jdz, jdz2, _, _ = f.dz_colored(w)
    betas = [k / 15.0 for k in range(16)]
    indptr, indices, vals = f.jbeta(jac, jdz, jdz2, betas)
    nnz = indices.numel()
    print(f"  nnz {nnz}, J(beta) values {vals.numel() * 16 / 1e9:.2f} GB", flush=True)
    assert vals.shape == (16, nnz)
    assert torch.all(vals[0].imag == 0)
    d = vals[15].imag
    for k in range(1, 15):
        assert torch.equal(vals[k].imag, betas[k] * d), k
    host = lambda t: t.cpu().numpy()
    mats = [tuple(map(host, f.csr(a))) for a in (jac, jdz, jdz2)]
    ip, ix = host(indptr), host(indices)
    vol = host(f.vol).reshape(case.ni, case.nj, order="F")
    rng = np.random.default_rng(5)
    n = 5 * case.im * case.jm
    rows = np.concatenate(([0, n - 1], rng.integers(0, n, 298)))
    for r in rows:
        parts = []
        for (p, c, v) in mats:
            parts.append(dict(zip(c[p[r]:p[r + 1]].tolist(), v[p[r]:p[r + 1]].tolist())))
        cols = sorted(set(parts[0]) | set(parts[1]) | set(parts[2]))
        np.testing.assert_array_equal(ix[ip[r]:ip[r + 1]], cols)
        vr = vol[r // (case.jm * 5) + 3, (r % (case.jm * 5)) // 5 + 3]
        jv = np.array([parts[0].get(c, 0.0) / vr if c in parts[0] else 0.0 for c in cols])
        dz = np.array([parts[1].get(c, 0.0) for c in cols])
        dzz = np.array([parts[2].get(c, 0.0) for c in cols])
        got = host(vals[:, ip[r]:ip[r + 1]])
        for k, b in enumerate(betas):
            np.testing.assert_array_equal(got[k].real, jv - (b * b) * dzz)
            np.testing.assert_array_equal(got[k].imag, b * dz)
