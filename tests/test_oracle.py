"""Pins the oracle (C++ restatement, oracle/bcoracle.cpp) before it is trusted.

* against the golden vectors written by the reference Fortran itself
  (tests/golden/make_golden.py);
* against what the mathematics fixes: a fluid at rest gives R = 0 on any mesh
  (every face flux reduces to p n and each closed cell sums its normals to 0),
  the tangent equals central finite differences of the residual, the tangent is
  linear in wd, and the colored 245-seed Jacobian equals the brute-force
  unit-seed Jacobian;
* against the reference library live, when it was built in this container;
* the adjoint (colored transposition of the tangent) against the reference's
  TAPENADE adjoint vectors, against brute-force unit seeds on every cell of a
  small array (which pins the dependence window it assumes), and by the
  dot-product test <v, J u> = <J^T v, u>.
"""
import os

import numpy as np
import scipy.sparse as sp
import pytest

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case, cylinder_case
from golden_util import load, case_from, close

CASES = ["bl24x16", "bl13x9", "cyl28x12"]


@pytest.fixture(scope="module")
def O():
    return oracle.Oracle()


@pytest.mark.parametrize("name", CASES)
def test_geometry_and_bcs_bit_exact(O, name):
    fx = load(name)
    case = case_from(fx)
    g = D.geometry(O, case)
    for k in ("nx", "ny", "xc", "yc", "vol", "volf"):
        np.testing.assert_array_equal(g[k], fx[f"geo_{k}"], err_msg=k)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    np.testing.assert_array_equal(w, fx["w_bc"])


@pytest.mark.parametrize("name", CASES)
def test_residual_vs_reference(O, name):
    fx = load(name)
    case = case_from(fx)
    g = {k: fx[f"geo_{k}"] for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    res = D.residual(O, case, g, np.asfortranarray(fx["w_bc"]))
    ok, err = close(res, fx["res"], 1e-13)
    assert ok, err
    norm, ninf = O.compute_norml2inf(res, case.im, case.jm, case.gh)
    np.testing.assert_allclose(norm, fx["norm"], rtol=1e-12)
    np.testing.assert_allclose(ninf, fx["ninf"], rtol=1e-12)


@pytest.mark.parametrize("name", CASES)
def test_tangent_vs_reference(O, name):
    fx = load(name)
    case = case_from(fx)
    g = {k: fx[f"geo_{k}"] for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    w = np.asfortranarray(fx["w_bc"])
    for s in range(2):
        wcopy = np.array(w, order="F", copy=True)
        d = np.array(fx["wd"][s], order="F", copy=True)
        D.apply_bcs(O, case, g, wcopy, wd=d)
        np.testing.assert_allclose(d, fx["wd_bc"][s], rtol=1e-13, atol=1e-13 * np.abs(d).max())
        _, resd = D.jvp(O, case, g, w, fx["wd"][s])
        ok, err = close(resd, fx["resd"][s], 1e-12)
        assert ok, err


@pytest.mark.parametrize("name", CASES)
def test_norml2_vs_reference_and_closed_form(O, name):
    """compute_norml2 (srcfv/norm.F90:2-32): per-variable L2 of the interior
    and nmoy = norm / (im jm).  Pinned by the reference's own outputs in the
    fixtures, by the numpy closed form on an array whose ghost frame is
    non-zero (so a sum that strays into the ghosts fails), and by the live
    reference on that array when it is built here."""
    fx = load(name)
    im, jm, gh = int(fx["im"]), int(fx["jm"]), int(fx["gh"])
    norm, nmoy = O.compute_norml2(np.asfortranarray(fx["res"]), im, jm, gh)
    np.testing.assert_allclose(norm, fx["norm"], rtol=1e-14)
    np.testing.assert_allclose(nmoy, fx["nmoy"], rtol=1e-14)
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.standard_normal(fx["res"].shape) * np.array([1.0, 10.0, 0.1, 1e-3, 100.0]))
    norm, nmoy = O.compute_norml2(a, im, jm, gh)
    inner = a[gh:-gh, gh:-gh, :]
    ref = np.sqrt((inner * inner).sum(axis=(0, 1)))
    np.testing.assert_allclose(norm, ref, rtol=1e-13)
    np.testing.assert_allclose(nmoy, ref / (im * jm), rtol=1e-13)
    if os.path.exists(oracle.REF_SO):
        rn, rm = oracle.Reference().compute_norml2(a, im, jm, gh)
        np.testing.assert_allclose(norm, rn, rtol=1e-14)
        np.testing.assert_allclose(nmoy, rm, rtol=1e-14)


@pytest.mark.parametrize("name", CASES)
def test_set_bndbl_vs_reference_and_closed_form(O, name):
    """set_bndbl_2d (set_bnd.f90:2-24), pure copies, bit-exact: against the
    reference's outputs in the fixtures and against the index formulas
    field(j, d, :) = w(1-d, j, :), wbd(i, :) = w(i-gh, jm, :) (Fortran
    indices, lower bound 1-gh) on a state with distinct values everywhere."""
    fx = load(name)
    im, jm, gh = int(fx["im"]), int(fx["jm"]), int(fx["gh"])
    rng = np.random.default_rng(31)
    field = np.asfortranarray(rng.standard_normal((jm, gh, 5)))
    wbd = np.asfortranarray(rng.standard_normal((im + gh, 5)))
    w = np.asfortranarray(fx["w_bc"])
    O.set_bndbl_2d(w, field, wbd, im, jm, gh)
    np.testing.assert_array_equal(field, fx["sbd_field"])
    np.testing.assert_array_equal(wbd, fx["sbd_wbd"])
    w = np.asfortranarray(rng.standard_normal(w.shape))
    field = np.zeros((jm, gh, 5), order="F")
    wbd = np.zeros((im + gh, 5), order="F")
    O.set_bndbl_2d(w, field, wbd, im, jm, gh)
    for d in range(1, gh + 1):
        np.testing.assert_array_equal(field[:, d - 1, :], w[gh - d, gh:gh + jm, :])
    np.testing.assert_array_equal(wbd, w[0:im + gh, gh + jm - 1, :])


def _norms_of_rest(O, case):
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    p = case.phys.p_inf
    w[:, :, 0] = 1.0
    w[:, :, 1:4] = 0.0
    w[:, :, 4] = p / (case.phys.gam - 1.0)
    wbd = np.zeros_like(case.wbd)
    wbd[:, 0] = 1.0
    wbd[:, 4] = p / (case.phys.gam - 1.0)
    case.wbd = np.asfortranarray(wbd)
    if case.field is not None:
        f = np.zeros_like(case.field)
        f[..., 0] = 1.0
        f[..., 4] = p / (case.phys.gam - 1.0)
        case.field = np.asfortranarray(f)
    D.apply_bcs(O, case, g, w)
    res = D.residual(O, case, g, w)
    # face fluxes are O(p |n|); the residual of a fluid at rest is pure rounding
    scale = p * np.abs(g["nx"]).max()
    return np.abs(res[3:-3, 3:-3, :]).max() / scale


# Only on the flat-plate block: on the O-mesh the reference's ghost j-face
# normals of column im+1 come from linearly extrapolated nodes (computegeom_2d
# then the partial periodic copy of cylinder.py:514-522), so the viscous
# contour of the last i-face does not close and a fluid at rest keeps an O(h^2)
# heat-flux residual there, in the reference as in the oracle.
@pytest.mark.parametrize("mk", [lambda: bl_case(20, 12, seed=0),
                                lambda: bl_case(9, 7, seed=0, ratio=1.2, dy1=1e-2)])
def test_fluid_at_rest_has_zero_residual(O, mk):
    assert _norms_of_rest(O, mk()) < 1e-13


@pytest.mark.parametrize("mk", [lambda: bl_case(16, 10, seed=7),
                                lambda: cylinder_case(14, 9, seed=8)])
def test_tangent_is_derivative(O, mk):
    case = mk()
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    rng = np.random.default_rng(3)
    wd = np.zeros_like(w, order="F")
    wd[3:-3, 3:-3, :] = rng.standard_normal((case.im, case.jm, 5)) * np.abs(w[3:-3, 3:-3, :])
    _, resd = D.jvp(O, case, g, w, wd)
    h = 1e-6
    rp, rm = [], []
    for sgn, out in ((1, rp), (-1, rm)):
        ws = np.array(w + sgn * h * wd, order="F")
        D.apply_bcs(O, case, g, ws)
        out.append(D.residual(O, case, g, ws))
    fd = (rp[0] - rm[0]) / (2 * h)
    ok, err = close(resd[3:-3, 3:-3], fd[3:-3, 3:-3], 1e-6)
    assert ok, err
    # linearity in wd
    _, r2 = D.jvp(O, case, g, w, 2.5 * wd)
    ok, err = close(r2, 2.5 * resd, 1e-14)
    assert ok, err


def test_colored_equals_bruteforce(O):
    case = bl_case(10, 8, seed=9)
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    Jc = D.colored_jacobian(O, case, g, w, prune=0.0).toarray()
    Jb = D.brute_jacobian(O, case, g, w)
    scale = np.abs(Jb).max(axis=1, keepdims=True)
    assert np.all(np.abs(Jc - Jb) <= 1e-13 * np.maximum(scale, 1e-300))


@pytest.mark.parametrize("name", ["jac_bl14x10", "jac_cyl14x8"])
def test_colored_jacobian_vs_reference(O, name):
    import scipy.sparse as sp

    fx = load(name)
    case = case_from(fx)
    g = D.geometry(O, case)
    w = np.asfortranarray(fx["w_bc"])
    J = D.colored_jacobian(O, case, g, w, coefdiag=fx["coefdiag"])
    n = 5 * case.im * case.jm
    Jref = sp.csr_matrix((fx["data"], fx["indices"], fx["indptr"]), shape=(n, n))
    diff = abs(J - Jref)
    rowscale = abs(Jref).max(axis=1).toarray().ravel()
    assert np.all(diff.max(axis=1).toarray().ravel() <= 1e-12 * rowscale)


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference library not built here")
def test_scatter_indices_match_reference(O):
    R = oracle.Reference()
    case = bl_case(15, 9, seed=2)
    im, jm, gh = case.im, case.jm, case.gh
    nb = im * jm * 49 * 25
    rng = np.random.default_rng(0)
    resd = np.asfortranarray(rng.standard_normal((case.ni, case.nj, 5)))
    cd = np.asfortranarray(rng.standard_normal((im, jm)))
    for withjn in (False, True):
        for coef in (None, cd):
            outs = []
            for lib in (O, R):
                jac, ia, ja = np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)
                for m, l, k in [(0, 0, 0), (2, 5, 3), (4, 6, 6), (1, 3, 0), (3, 6, 2)]:
                    lib.computejacobianfromjv(jac, ia, ja, resd, m, l, k, gh, im, jm,
                                              coefdiag=coef, withjn=withjn)
                outs.append((jac, ia, ja))
            for a, b in zip(*outs):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", CASES)
def test_adjoint_vs_reference(O, name):
    fx = load(name)
    case = case_from(fx)
    g = {k: fx[f"geo_{k}"] for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    wb = D.vjp(O, case, g, fx["w_bc"], fx["resb"])
    ok, err = close(wb, fx["wb"], 1e-12)
    assert ok, err


def _unit_seed_vjp(O, case, g, w, resb):
    n = case.ni * case.nj * 5
    wb = np.zeros(n)
    for c in range(n):
        wd = np.zeros(n)
        wd[c] = 1.0
        resd = D.jvp_flux(O, case, g, w, wd.reshape(w.shape, order="F"))
        wb[c] = np.sum(resd * resb)
    return wb.reshape(w.shape, order="F")


@pytest.mark.parametrize("mk", [lambda: bl_case(9, 7, seed=4), lambda: cylinder_case(10, 6, seed=5)])
def test_adjoint_equals_unit_seed_transpose(O, mk):
    case = mk()
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    rng = np.random.default_rng(11)
    resb = np.zeros_like(w, order="F")
    resb[3:-3, 3:-3, :] = rng.standard_normal((case.im, case.jm, 5))
    wb = D.vjp(O, case, g, w, resb)
    ref = _unit_seed_vjp(O, case, g, w, resb)
    ok, err = close(wb, ref, 1e-13)
    assert ok, err
    # every ghost-frame cell the residual reads gets an adjoint; corners do not
    assert np.abs(wb[:3, 3:-3]).max() > 0 and np.abs(wb[3:-3, :3]).max() > 0


def test_adjoint_dot_product(O):
    case = bl_case(18, 11, seed=12)
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    rng = np.random.default_rng(13)
    u = np.asfortranarray(rng.standard_normal(w.shape) * np.abs(w))
    v = np.zeros_like(w, order="F")
    v[3:-3, 3:-3, :] = rng.standard_normal((case.im, case.jm, 5))
    Ju = D.jvp_flux(O, case, g, w, u)
    JTv = D.vjp(O, case, g, w, v)
    a, b = np.sum(v * Ju), np.sum(JTv * u)
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_two_zone_colored_equals_bruteforce(O):
    """O-mesh with im not a multiple of 7: the two-zone coloring of
    cylinder.py:895-935 recovers the true Jacobian across the periodic seam
    (a single 7-periodic coloring would alias the wrap columns)."""
    case = cylinder_case(17, 8, seed=14)
    assert D.two_zone_split(case) == 7
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    Jc = D.colored_jacobian(O, case, g, w, prune=0.0).toarray()
    Jb = D.brute_jacobian(O, case, g, w)
    scale = np.abs(Jb).max(axis=1, keepdims=True)
    assert np.all(np.abs(Jc - Jb) <= 1e-13 * np.maximum(scale, 1e-300))


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference library not built here")
def test_two_zone_seeds_and_scatter_match_reference(O):
    R = oracle.Reference()
    im, jm, gh = 17, 9, 3
    ni, nj = im + 2 * gh, jm + 2 * gh
    nb = im * jm * 49 * 25 * 2
    rng = np.random.default_rng(4)
    resd = np.asfortranarray(rng.standard_normal((ni, nj, 5)))
    cd = np.asfortranarray(rng.standard_normal((im, jm)))
    for zone in ((0, 6, 0, jm - 1), (7, im - 1, 0, jm - 1), (3, 12, 2, 6)):
        a, b = np.zeros((ni, nj, 5), order="F"), np.zeros((ni, nj, 5), order="F")
        O.testvector_partial(a, 2, 4, 1, gh, im, jm, *zone)
        R.testvector_partial(b, 2, 4, 1, gh, im, jm, *zone)
        np.testing.assert_array_equal(a, b)
    # slots prefilled with a mix of zeros and values exercise the occupied-slot
    # branches of the check
    pre = np.where(rng.random(nb) < 0.3, rng.standard_normal(nb), 0.0)
    outs = []
    for lib in (O, R):
        jac, ia, ja = pre.copy(), np.full(nb, 7, np.int32), np.full(nb, 9, np.int32)
        for n in (0, 1):
            for m, l, k in [(0, 0, 0), (2, 5, 3), (4, 6, 6), (1, 3, 0), (3, 2, 5)]:
                lib.computejacobianfromjv_relaxed_withjnandcheck(jac, ia, ja, resd, m, l, k, gh, im, jm, cd,
                                                                 2e-16, n)
        outs.append((jac, ia, ja))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("im", [17, 36, 64])
def test_two_zone_relaxation_counts_once_per_zone(O, im):
    """Reference quirk of the two-zone loop (cylinder.py:905-933): each zone's
    relaxed scatter writes coefdiag - resd into its own slot whenever the
    slot's column is the row's own cell, so the summed CSR carries coefdiag
    once per zone whose window holds the diagonal -- twice on most rows.
    With resd = 0 every kept entry is such a diagonal slot: the scatter of
    all 490 seeds (the reference build where present, and the restatement)
    gives diag(mult * coefdiag), mult from the window rules."""
    from golden_util import two_zone_diag_multiplicity

    gh, jm = 3, 9
    ni, nj = im + 2 * gh, jm + 2 * gh
    nb = im * jm * 49 * 25 * 2
    split = (im // 2) // 7 * 7
    cd = np.asfortranarray(np.random.default_rng(5).uniform(0.5, 1.5, (im, jm)))
    resd = np.zeros((ni, nj, 5), order="F")
    libs = [O] + ([oracle.Reference()] if os.path.exists(oracle.REF_SO) else [])
    mult = two_zone_diag_multiplicity(im, gh)
    assert set(mult) <= {1, 2} and (mult == 2).any()
    want = np.repeat((mult[:, None] * cd)[:, :, None], 5, axis=2).ravel()  # rows e + 5 (j + jm i)
    for lib in libs:
        jac, ia, ja = np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        for n in (0, 1):
            for m in range(5):
                for l in range(7):
                    for k in range(7):
                        lib.computejacobianfromjv_relaxed_withjnandcheck(jac, ia, ja, resd, m, l, k, gh, im, jm,
                                                                         cd, 2e-16, n)
        keep = np.abs(jac) > 2e-16
        A = sp.csr_matrix((jac[keep], (ia[keep], ja[keep])), shape=(5 * im * jm, 5 * im * jm))
        assert A.nnz == 5 * im * jm and np.all(A.diagonal() != 0.0)
        np.testing.assert_allclose(A.diagonal(), want, rtol=1e-15)
    assert split >= 7


# ------------------------------------------------------------------ Dz, Dzz
def _uniform(case, U=0.3):
    """rho = 1, u = U, v = w = 0 at the free-stream pressure everywhere."""
    w = np.zeros((case.ni, case.nj, 5), order="F")
    p = case.phys.p_inf
    w[:, :, 0] = 1.0
    w[:, :, 1] = U
    w[:, :, 4] = p / (case.phys.gam - 1.0) + 0.5 * U * U
    return w


@pytest.mark.parametrize("name", CASES)
def test_dz_vs_reference(O, name):
    fx = load(name)
    case = case_from(fx)
    g = {k: np.asfortranarray(fx[f"geo_{k}"]) for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    a = case.phys.scheme_args()[:9]
    for fn, key in ((O.coeffs_5p_dz, "dz"), (O.coeffs_5p_dz2, "dz2")):
        out = np.zeros_like(fx["w_bc"], order="F")
        fn(out, np.asfortranarray(fx["w_bc"]), np.asfortranarray(fx["wd_bc"][0]), g["x0"], g["y0"], g["nx"],
           g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], case.gh, *a, case.im, case.jm)
        ok, err = close(out, fx[key], 1e-13)
        assert ok, (key, err)


def test_dz_closed_forms(O):
    """Uniform rho = 1, u = U state: Dz of a unit spanwise-momentum tangent is
    the z-flux Jacobian column (1, u, v, 0, H); Dzz of a unit streamwise
    momentum tangent is (0, -mu, 0, 0, mu U (gamma/Pr - 1)) (viscous and
    conductive z-diffusion of u and of T = e/cv)."""
    case = bl_case(12, 9, seed=0, ratio=1.1, dy1=1e-2)
    g = D.geometry(O, case)
    U = 0.3
    w = _uniform(case, U)
    a = case.phys.scheme_args()
    ph = case.phys
    e4 = np.zeros_like(w, order="F")
    e4[:, :, 3] = 1.0
    dz = np.zeros_like(w, order="F")
    O.coeffs_5p_dz(dz, w, e4, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], 3,
                   *a[:9], case.im, case.jm)
    H = w[0, 0, 4] + ph.p_inf
    np.testing.assert_allclose(dz[3:-3, 3:-3].reshape(-1, 5), np.tile([1.0, U, 0.0, 0.0, H], (case.im * case.jm, 1)),
                               rtol=1e-14, atol=1e-14)
    e2 = np.zeros_like(w, order="F")
    e2[:, :, 1] = 1.0
    dz2 = np.zeros_like(w, order="F")
    O.coeffs_5p_dz2(dz2, w, e2, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], 3,
                    *a[:9], case.im, case.jm)
    cp, cv, pr = a[0], a[1], a[2]
    t = (w[0, 0, 4] - 0.5 * U * U) / cv
    betas = ph.muref * (ph.tref + ph.cs) / (np.sqrt(ph.tref) * ph.tref)
    mu = betas / (t + ph.cs) * np.sqrt(t) * t
    ref = np.array([0.0, -mu, 0.0, 0.0, mu * U * (cp / cv / pr - 1.0)])
    np.testing.assert_allclose(dz2[3:-3, 3:-3].reshape(-1, 5), np.tile(ref, (case.im * case.jm, 1)), rtol=1e-13,
                               atol=1e-15)


def test_dz_stencil_and_linearity(O):
    """Dzz is pointwise (block diagonal); Dz at a cell reads only the cross of
    cells (i +- 1, 2; j +- 1, 2); both are linear in wd."""
    case = bl_case(14, 10, seed=5)
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    a = case.phys.scheme_args()[:9]
    rng = np.random.default_rng(8)

    def ops(wd):
        out = []
        for fn in (O.coeffs_5p_dz, O.coeffs_5p_dz2):
            r = np.zeros_like(w, order="F")
            fn(r, w, np.asfortranarray(wd), g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"],
               g["volf"], 3, *a, case.im, case.jm)
            out.append(r)
        return out

    wd = np.asfortranarray(rng.standard_normal(w.shape))
    dz, dz2 = ops(wd)
    ci, cj = 3 + 6, 3 + 5  # interior cell (7, 6) in storage coordinates
    for (di, dj, cross) in ((1, 0, True), (0, -2, True), (1, 1, False), (3, 0, False), (0, 3, False)):
        p = np.array(wd, order="F", copy=True)
        p[ci + di, cj + dj, :] += rng.standard_normal(5)
        pz, pz2 = ops(p)
        assert np.array_equal(pz2[ci, cj], dz2[ci, cj])
        assert np.array_equal(pz[ci, cj], dz[ci, cj]) != cross
    dzs, dz2s = ops(2.5 * wd)
    ok, err = close(dzs, 2.5 * dz, 1e-14)
    assert ok, err
    ok, err = close(dz2s, 2.5 * dz2, 1e-14)
    assert ok, err


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="reference library not built here")
def test_dz_scatter_matches_reference(O):
    R = oracle.Reference()
    im, jm, gh = 15, 11, 3
    nb = im * jm * 49 * 25
    rng = np.random.default_rng(2)
    dz = np.asfortranarray(rng.standard_normal((im + 6, jm + 6, 5)))
    outs = []
    for lib in (O, R):
        jac, ia, ja = np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        for m, l, k in [(0, 0, 0), (2, 5, 3), (4, 6, 6), (1, 3, 0), (3, 2, 5), (0, 4, 4)]:
            lib.computejacobianfromdz(jac, ia, ja, dz, m, l, k, gh, im, jm)
        outs.append((jac, ia, ja))
    for x, y in zip(*outs):
        np.testing.assert_array_equal(x, y)


def _bc_calls(O, case, g):
    im, jm, gh, gam, it = case.im, case.jm, case.gh, case.phys.gam, case.interf
    calls = {
        "Jhi": lambda a, ad: O.bc_no_reflexion_2d(a, case.wbd, "Jhi", it["Jhi"], g["nx"], g["ny"], gam, gh, im, jm,
                                                  wd=ad),
        "Jlo": lambda a, ad: O.bc_wall_viscous_adia_2d(a, "Jlo", gam, it["Jlo"], gh, im, jm, wd=ad),
    }
    if case.kind == "bl":
        calls["Ilo"] = lambda a, ad: O.bc_supandsubinlet_2d(a, "Ilo", it["Ilo"], case.field, g["nx"], g["ny"], gam,
                                                            im, jm, gh, wd=ad)
        calls["Ihi"] = lambda a, ad: O.bc_extrapolate_o2_2d(a, "Ihi", it["Ihi"], im, jm, gh, wd=ad)
    return calls


@pytest.mark.parametrize("name", CASES)
def test_bc_adjoints_vs_reference(O, name):
    """Transpose of the restated tangent BCs against the reference's TAPENADE
    adjoints bc_*_b (golden vectors): each BC alone on one seeded wb."""
    fx = load(name)
    case = case_from(fx)
    g = {k: fx[f"geo_{k}"] for k in ("nx", "ny")}
    w = np.asfortranarray(fx["w_bc"])
    for loc, fn in _bc_calls(O, case, g).items():
        wb = D.adjoint_line_bc(case, fn, w, fx["bcb_wb"], loc, case.interf[loc])
        ok, err = close(wb, fx[f"bcb_{loc}"], 1e-13)
        assert ok, (loc, err)


@pytest.mark.parametrize("mk", [lambda: bl_case(16, 10, seed=31), lambda: cylinder_case(22, 9, seed=32)])
def test_bc_sequence_adjoint_dot_product(O, mk):
    """<v, BCd(u)> = <BCb(v), u> for the whole BC sequence (joins included):
    BCd replaces the ghost frame of u by the linearised BCs, BCb is the
    sequence's adjoint in reverse order."""
    case = mk()
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    rng = np.random.default_rng(33)
    u = np.asfortranarray(rng.standard_normal(w.shape))
    v = np.asfortranarray(rng.standard_normal(w.shape))
    wc, ud = np.array(w, order="F", copy=True), np.array(u, order="F", copy=True)
    D.apply_bcs(O, case, g, wc, wd=ud)
    vb = D.adjoint_bcs(O, case, g, w, v)
    a, b = np.sum(v * ud), np.sum(vb * u)
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))
    # the ghost adjoints the sequence writes are consumed
    assert np.abs(vb[:, :3]).max() == 0.0 and np.abs(vb[:, -3:]).max() == 0.0
