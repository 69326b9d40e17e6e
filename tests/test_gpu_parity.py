"""HIP path (through the C ABI) against the reference's golden vectors and the
oracle, on the same seeded inputs.

Tolerances (FP64 everywhere): the scheme's residual is a difference of O(1)
face fluxes, so pointwise parity is stated normwise per variable:
max|a - b| <= tol * max|b| over the interior, tol = 1e-12 for residual, JVP
and Jacobian rows (the project's parity bound), 1e-14 for geometry and
boundary-condition ghost states.  Integer outputs (COO indices) are bit-exact.
"""
import numpy as np
import pytest

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case, cylinder_case
from golden_util import load, case_from, close

pytestmark = pytest.mark.gpu

CASES = ["bl24x16", "bl13x9", "cyl28x12"]
GEO = ("nx", "ny", "xc", "yc", "vol", "volf")


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


@pytest.fixture(scope="module")
def O():
    return oracle.Oracle()


def _flow(case):
    from broadcast_amd.flow import Flow

    return Flow(case)


def _host(f, t, shape):
    from broadcast_amd.flow import to_host

    return to_host(t, shape)


@pytest.mark.parametrize("name", CASES)
def test_geometry_matches_reference(torch_mod, name):
    fx = load(name)
    f = _flow(case_from(fx))
    ni, nj = f.ni, f.nj
    shapes = dict(nx=(ni + 1, nj + 1, 2), ny=(ni + 1, nj + 1, 2), xc=(ni, nj), yc=(ni, nj),
                  vol=(ni, nj), volf=(ni, nj, 2))
    for k in GEO:
        got = _host(f, getattr(f, k), shapes[k])
        ref = fx[f"geo_{k}"]
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max(), err_msg=k)


@pytest.mark.parametrize("name", CASES)
def test_bcs_residual_jvp_match_reference(torch_mod, name):
    from broadcast_amd.flow import to_device

    fx = load(name)
    f = _flow(case_from(fx))
    w = to_device(fx["w_in"])
    f.apply_bcs(w)
    wbc = _host(f, w, f.shape_w)
    ok, err = close(wbc, fx["w_bc"], 1e-14)
    assert ok, f"ghost states rel err per variable {err}"
    res = _host(f, f.residual(w), f.shape_w)[3:-3, 3:-3]
    ok, err = close(res, fx["res"][3:-3, 3:-3], 1e-12)
    assert ok, f"residual rel err per variable {err}"
    norm, ninf = f.norms(to_device(fx["res"]))
    np.testing.assert_allclose(norm, fx["norm"], rtol=1e-13)
    np.testing.assert_allclose(ninf, fx["ninf"], rtol=1e-12)
    for s in range(2):
        wd = to_device(fx["wd"][s])
        _, resd = f.jvp(to_device(fx["w_bc"]), wd)
        ok, err = close(_host(f, wd, f.shape_w), fx["wd_bc"][s], 1e-13)
        assert ok, f"ghost tangents rel err per variable {err}"
        ok, err = close(_host(f, resd, f.shape_w)[3:-3, 3:-3], fx["resd"][s][3:-3, 3:-3], 1e-12)
        assert ok, f"jvp rel err per variable {err}"


@pytest.mark.parametrize("name", ["bl24x16", "cyl28x12"])
def test_dropin_modules_numpy_path(torch_mod, name):
    """The f2py call forms of BROADCAST_npz.py:1018-1035,1079-1085 on numpy
    F-order arrays, mutated in place."""
    from broadcast_amd.srcfv import f_bnd, f_lin, f_sch, f_norm, f_geom

    fx = load(name)
    case = case_from(fx)
    im, jm, gh = case.im, case.jm, case.gh
    g = {k: np.zeros_like(fx[f"geo_{k}"], order="F") for k in GEO}
    x0, y0 = np.array(fx["x0_in"], order="F"), np.array(fx["y0_in"], order="F")
    f_geom.computegeom_2d(x0, y0, g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], im, jm, gh)
    if case.kind == "bl":
        for k in GEO:
            np.testing.assert_allclose(g[k], fx[f"geo_{k}"], rtol=1e-14,
                                       atol=1e-14 * np.abs(fx[f"geo_{k}"]).max())
    geo = {k: np.array(fx[f"geo_{k}"], order="F") for k in GEO}
    it = case.interf
    gam = case.phys.gam
    w = np.array(fx["w_in"], order="F")
    if case.kind == "bl":
        f_bnd.bc_supandsubinlet_2d(w, "Ilo", it["Ilo"], case.field, geo["nx"], geo["ny"], gam, im, jm)
        f_bnd.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_bnd.bc_extrapolate_o2_2d(w, "Ihi", it["Ihi"], im, jm, gh)
        f_bnd.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm)
    else:
        from broadcast_amd.cases import join_boxes

        f_bnd.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_bnd.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm)
        (prr1, prd2, tr), (prr2, prd1, _) = join_boxes(im, jm, gh)
        f_bnd.jn_match_2d(w, prr1, gh, gh, gh, gh, im, jm, w, prd2, gh, gh, gh, gh, im, jm, tr)
        f_bnd.jn_match_2d(w, prr2, gh, gh, gh, gh, im, jm, w, prd1, gh, gh, gh, gh, im, jm, tr)
    ok, err = close(w, fx["w_bc"], 1e-14)
    assert ok, err
    res = np.zeros_like(w, order="F")
    args = case.phys.scheme_args()
    f_sch.flux_num_dnc5_2d(res, w, fx["geo_x0"], fx["geo_y0"], geo["nx"], geo["ny"], geo["xc"],
                           geo["yc"], geo["vol"], geo["volf"], gh, *args, im, jm)
    ok, err = close(res[3:-3, 3:-3], fx["res"][3:-3, 3:-3], 1e-12)
    assert ok, err
    norm, ninf = f_norm.compute_norml2inf(res, im, jm, gh)
    np.testing.assert_allclose(norm, fx["norm"], rtol=1e-11)
    if case.kind == "bl":
        wd = np.array(fx["wd"][0], order="F")
        f_lin.bc_supandsubinlet_2d_d(w, wd, "Ilo", it["Ilo"], case.field, geo["nx"], geo["ny"], gam, im, jm)
        f_lin.bc_no_reflexion_2d_d(w, wd, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_lin.bc_extrapolate_o2_2d_d(w, wd, "Ihi", it["Ihi"], im, jm, gh)
        f_lin.bc_wall_viscous_adia_2d_d(w, wd, "Jlo", gam, it["Jlo"], gh, im, jm)
        ok, err = close(wd, fx["wd_bc"][0], 1e-13)
        assert ok, err
        resd = np.zeros_like(w, order="F")
        res_before = res.copy(order="F")
        f_lin.flux_num_dnc5_2d_d(res, resd, w, wd, fx["geo_x0"], fx["geo_y0"], geo["nx"], geo["ny"],
                                 geo["xc"], geo["yc"], geo["vol"], geo["volf"], gh, *args, im, jm)
        ok, err = close(resd[3:-3, 3:-3], fx["resd"][0][3:-3, 3:-3], 1e-12)
        assert ok, err
        # the reference's tangent assigns residud only
        assert np.array_equal(res, res_before)


@pytest.mark.parametrize("name", CASES)
def test_nowall_variant_matches_reference(torch_mod, name):
    """flux_num_dnc5_nowall_2d and its tangent (srcfv/rhs/
    flux_num_dnc5_nowall.F90:7-176, srcfv/tangent/flux_num_dnc5_nowall_d.f90:
    15) through the drop-in f_sch / f_lin call forms, on the stored state after
    BCs and the first linearised direction.  The tangent leaves residu as
    given, as the reference's does."""
    from broadcast_amd.srcfv import f_lin, f_sch

    fx = load(name)
    case = case_from(fx)
    im, jm, gh = case.im, case.jm, case.gh
    geo = {k: np.array(fx[f"geo_{k}"], order="F") for k in ("x0", "y0") + GEO}
    args = case.phys.scheme_args()
    w = np.array(fx["w_bc"], order="F")
    res = np.zeros_like(w, order="F")
    f_sch.flux_num_dnc5_nowall_2d(res, w, geo["x0"], geo["y0"], geo["nx"], geo["ny"], geo["xc"], geo["yc"],
                                  geo["vol"], geo["volf"], gh, *args, im, jm)
    ok, err = close(res[3:-3, 3:-3], fx["res_nowall"][3:-3, 3:-3], 1e-12)
    assert ok, err
    # the variant differs from the wall scheme on the rows next to j = 1
    assert not np.allclose(res[3:-3, 3:5], fx["res"][3:-3, 3:5], rtol=1e-6, atol=0)
    wd = np.array(fx["wd_bc"][0], order="F")
    res2 = np.full_like(w, 7.0, order="F")
    resd = np.zeros_like(w, order="F")
    f_lin.flux_num_dnc5_nowall_2d_d(res2, resd, w, wd, geo["x0"], geo["y0"], geo["nx"], geo["ny"], geo["xc"],
                                    geo["yc"], geo["vol"], geo["volf"], gh, *args, im, jm)
    ok, err = close(resd[3:-3, 3:-3], fx["resd_nowall"][3:-3, 3:-3], 1e-12)
    assert ok, err
    assert (res2 == 7.0).all()


@pytest.mark.parametrize("name,order", [("dnc3_bl20x12", 3), ("dnc3_cyl21x10", 3), ("dnc7_bl22x14", 7),
                                        ("dnc7_cyl24x12", 7), ("dnc9_bl23x16", 9), ("dnc9_cyl26x13", 9)])
def test_other_orders_match_reference(torch_mod, name, order):
    """The schemes of order 3, 7 and 9 on arrays of gh = (order + 1) / 2,
    through the drop-in call forms: geometry, the case's BCs and their
    linearisation, flux_num_dnc{order}_2d and flux_num_dnc{order}_nowall_2d
    (srcfv/rhs) and their tangents (srcfv/tangent/flux_num_dnc{order}*_d.f90)
    against the reference's outputs."""
    from broadcast_amd.srcfv import f_bnd, f_geom, f_lin, f_sch

    fx = load(name)
    case = case_from(fx)
    im, jm, gh = case.im, case.jm, case.gh
    assert gh == (order + 1) // 2
    g = {k: np.zeros_like(fx[f"geo_{k}"], order="F") for k in GEO}
    x0, y0 = np.array(fx["x0_in"], order="F"), np.array(fx["y0_in"], order="F")
    f_geom.computegeom_2d(x0, y0, g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], im, jm, gh)
    if case.kind == "bl":
        # the O-mesh's stored geometry takes periodic ghost metrics after
        # computegeom (cylinder.py:499-527), which this call alone does not
        for k in ("nx", "ny", "vol", "volf"):
            ok, err = close(g[k], fx[f"geo_{k}"], 1e-14)
            assert ok, (k, err)
    geo = {k: np.array(fx[f"geo_{k}"], order="F") for k in ("x0", "y0") + GEO}
    it = case.interf
    gam = case.phys.gam
    w = np.array(fx["w_in"], order="F")
    wd = np.array(fx["wd"], order="F")
    if case.kind == "bl":
        f_bnd.bc_supandsubinlet_2d(w, "Ilo", it["Ilo"], case.field, geo["nx"], geo["ny"], gam, im, jm)
        f_bnd.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_bnd.bc_extrapolate_o2_2d(w, "Ihi", it["Ihi"], im, jm, gh)
        f_bnd.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm)
        w2 = np.array(w, order="F")
        f_lin.bc_supandsubinlet_2d_d(w2, wd, "Ilo", it["Ilo"], case.field, geo["nx"], geo["ny"], gam, im, jm)
        f_lin.bc_no_reflexion_2d_d(w2, wd, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_lin.bc_extrapolate_o2_2d_d(w2, wd, "Ihi", it["Ihi"], im, jm, gh)
        f_lin.bc_wall_viscous_adia_2d_d(w2, wd, "Jlo", gam, it["Jlo"], gh, im, jm)
        ok, err = close(wd, fx["wd_bc"], 1e-13)
        assert ok, err
    else:
        from broadcast_amd.cases import join_boxes

        f_bnd.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], geo["nx"], geo["ny"], gam, gh, im, jm)
        f_bnd.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm)
        (prr1, prd2, tr), (prr2, prd1, _) = join_boxes(im, jm, gh)
        for _ in range(2):
            f_bnd.jn_match_2d(w, prr1, gh, gh, gh, gh, im, jm, w, prd2, gh, gh, gh, gh, im, jm, tr)
            f_bnd.jn_match_2d(w, prr2, gh, gh, gh, gh, im, jm, w, prd1, gh, gh, gh, gh, im, jm, tr)
        wd = np.array(fx["wd_bc"], order="F")
    ok, err = close(w, fx["w_bc"], 1e-14)
    assert ok, err
    args = case.phys.scheme_args()
    common = (geo["x0"], geo["y0"], geo["nx"], geo["ny"], geo["xc"], geo["yc"], geo["vol"], geo["volf"], gh,
              *args, im, jm)
    sl = (slice(gh, -gh), slice(gh, -gh))
    for tag in ("", "_nowall"):
        fn = getattr(f_sch, f"flux_num_dnc{order}{tag}_2d")
        fnd = getattr(f_lin, f"flux_num_dnc{order}{tag}_2d_d")
        res = np.zeros_like(w, order="F")
        fn(res, w, *common)
        ok, err = close(res[sl], fx[f"res{order}{tag}"][sl], 1e-12)
        assert ok, (tag, err)
        res2 = np.full_like(w, 7.0, order="F")
        resd = np.zeros_like(w, order="F")
        fnd(res2, resd, w, np.array(fx["wd_bc"], order="F"), *common)
        ok, err = close(resd[sl], fx[f"resd{order}{tag}"][sl], 1e-12)
        assert ok, (tag, err)
        assert (res2 == 7.0).all()


def test_scheme_orders_refuse_wrong_ghost_width(torch_mod):
    """The 5th-order entries and the Jacobian paths need gh = 3; on a gh = 2
    context they fail with a message instead of reading past the arrays."""
    import ctypes as C

    from broadcast_amd import _lib
    from broadcast_amd import device as dev

    torch = torch_mod
    ctx = dev.context(12, 10, 2)
    L = _lib.lib()
    n = (12 + 4) * (10 + 4) * 5
    a = torch.zeros(n, dtype=torch.float64, device="cuda")
    nn = (12 + 5) * (10 + 5) * 2
    b = torch.ones(nn, dtype=torch.float64, device="cuda")
    g = _lib.Metrics(b.data_ptr(), b.data_ptr(), a.data_ptr(), a.data_ptr())
    ph = dev.phys(1004.5, 717.5, 0.72, 1.4, 287.0, 110.4, 1.7e-5, 273.0, 110.4, 0.0, 1.0)
    rc = L.bca_residual_dnc5(ctx.h, a.data_ptr(), a.data_ptr(), g, ph)
    assert rc != 0 and b"gh = 3" in L.bca_last_error()


@pytest.mark.parametrize("name", ["jac_bl14x10", "jac_cyl14x8"])
def test_colored_jacobian_matches_reference(torch_mod, name):
    import scipy.sparse as sp
    from broadcast_amd.flow import to_device

    fx = load(name)
    case = case_from(fx)
    f = _flow(case)
    w = to_device(fx["w_bc"])
    jac, ia, ja = f.jacobian_colored(w, coefdiag=to_device(fx["coefdiag"]))
    jac, ia, ja = jac.cpu().numpy(), ia.cpu().numpy(), ja.cpu().numpy()
    keep = np.abs(jac) > 2e-16  # remove_zero_jac (BROADCAST_npz.py:129-135)
    n = 5 * case.im * case.jm
    J = sp.csr_matrix((jac[keep], (ia[keep], ja[keep])), shape=(n, n))
    Jref = sp.csr_matrix((fx["data"], fx["indices"], fx["indptr"]), shape=(n, n))
    diff = abs(J - Jref).max(axis=1).toarray().ravel()
    rowscale = abs(Jref).max(axis=1).toarray().ravel()
    assert np.all(diff <= 1e-12 * rowscale), np.max(diff / rowscale)
    # the same through the device prune + CSR
    indptr, indices, data = f.csr(to_device(jac))
    Jg = sp.csr_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
    np.testing.assert_array_equal(Jg.indptr, J.indptr)
    np.testing.assert_array_equal(Jg.indices, J.indices)
    np.testing.assert_array_equal(Jg.data, J.data)


def test_seed_and_scatter_match_oracle(torch_mod, O):
    """testvector + computejacobianfromjv[_relaxed][_withjn] through f_misc,
    indices bit-exact, values exact (pure copies/negations)."""
    from broadcast_amd.misc import f_misc

    case = bl_case(17, 11, seed=3)
    im, jm, gh = case.im, case.jm, case.gh
    rng = np.random.default_rng(5)
    resd = np.asfortranarray(rng.standard_normal((case.ni, case.nj, 5)))
    cd = np.asfortranarray(rng.standard_normal((im, jm)))
    nb = im * jm * 49 * 25
    for (m, l, k) in [(0, 0, 0), (3, 6, 2), (4, 2, 6), (1, 5, 5)]:
        wd_g = np.zeros((case.ni, case.nj, 5), order="F")
        wd_o = np.zeros_like(wd_g)
        f_misc.testvector(wd_g, m, l, k, gh, im, jm)
        O.testvector(wd_o, m, l, k, gh, im, jm)
        np.testing.assert_array_equal(wd_g, wd_o)
        for coef, withjn, fn in ((None, False, "plain"), (cd, False, "relaxed"), (cd, True, "relaxed_withjn"),
                                 (None, True, "withjn")):
            g = [np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)]
            o = [np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)]
            if fn == "plain":
                f_misc.computejacobianfromjv(*g, resd, m, l, k, gh, im, jm)
            elif fn == "relaxed":
                f_misc.computejacobianfromjv_relaxed(*g, resd, m, l, k, gh, cd)
            elif fn == "relaxed_withjn":
                f_misc.computejacobianfromjv_relaxed_withjn(*g, resd, m, l, k, gh, cd)
            else:
                f_misc.computejacobianfromjv_withjn(*g, resd, m, l, k, gh, im, jm)
            O.computejacobianfromjv(*o, resd, m, l, k, gh, im, jm, coefdiag=coef, withjn=withjn)
            for a, b in zip(g, o):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mk", [lambda: bl_case(96, 64, seed=21, reynolds=3e5),
                                lambda: cylinder_case(70, 48, seed=22)])
def test_larger_meshes_match_oracle(torch_mod, O, mk):
    from broadcast_amd.flow import to_device

    case = mk()
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    res_o = D.residual(O, case, g, w)
    rng = np.random.default_rng(9)
    wd = np.zeros_like(w, order="F")
    wd[3:-3, 3:-3, :] = rng.standard_normal((case.im, case.jm, 5))
    _, resd_o = D.jvp(O, case, g, w, wd)
    f = _flow(case)
    wg = to_device(case.w)
    f.apply_bcs(wg)
    ok, err = close(_host(f, wg, f.shape_w), w, 1e-14)
    assert ok, err
    ok, err = close(_host(f, f.residual(wg), f.shape_w)[3:-3, 3:-3], res_o[3:-3, 3:-3], 1e-12)
    assert ok, err
    _, resd = f.jvp(wg, to_device(wd))
    ok, err = close(_host(f, resd, f.shape_w)[3:-3, 3:-3], resd_o[3:-3, 3:-3], 1e-12)
    assert ok, err


@pytest.mark.parametrize("mk", [lambda: bl_case(29, 18, seed=31),
                                lambda: cylinder_case(28, 14, seed=32),
                                lambda: cylinder_case(30, 12, seed=33)])
def test_batched_jacobian_coo_matches_oracle_loop(torch_mod, O, mk):
    """The whole COO of the batched assembly against the oracle running the
    reference's seed loop (testvector + tangent BCs + JVP + scatter; for the
    O-mesh with im = 30, not a multiple of 7, the two-zone loop of
    testvector_partial + computejacobianfromjv_relaxed_withjnandcheck):
    indices bit-exact, values to 1e-12 of each row's largest entry."""
    from broadcast_amd.flow import to_device

    case = mk()
    print(f"case {case.kind} {case.im}x{case.jm} two-zone split {D.two_zone_split(case)}", flush=True)
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    im, jm = case.im, case.jm
    rng = np.random.default_rng(4)
    cd = np.asfortranarray(rng.uniform(0.5, 1.5, (im, jm)))
    jac_o, ia_o, ja_o = D.colored_jacobian(O, case, g, w, coefdiag=cd, coo=True)
    f = _flow(case)
    assert f.nbentry == jac_o.shape[0]
    wg = to_device(case.w)
    f.apply_bcs(wg)
    jac, ia, ja = f.jacobian_colored(wg, coefdiag=to_device(cd))
    ia, ja, jac = ia.cpu().numpy(), ja.cpu().numpy(), jac.cpu().numpy()
    np.testing.assert_array_equal(ia, ia_o)
    np.testing.assert_array_equal(ja, ja_o)
    n = 5 * im * jm
    rowmax = np.zeros(n)
    np.maximum.at(rowmax, ia_o, np.abs(jac_o))
    err = np.abs(jac - jac_o)
    bad = err > 1e-12 * np.maximum(rowmax[ia_o], 1e-300)
    assert not bad.any(), (int(bad.sum()), float((err / np.maximum(rowmax[ia_o], 1e-300)).max()))
    # device prune + CSR: bit-exact against scipy on the oracle's COO values
    # (remove_zero_jac + csr_matrix, BROADCAST_npz.py:129-135,
    # misc/PETSc_func.py:85), and end to end within the Jacobian tolerance
    import scipy.sparse as sp

    keep = np.abs(jac_o) > 2e-16
    Jo = sp.csr_matrix((jac_o[keep], (ia_o[keep], ja_o[keep])), shape=(n, n))
    indptr, indices, data = f.csr(to_device(jac_o))
    np.testing.assert_array_equal(indptr.cpu().numpy(), Jo.indptr)
    np.testing.assert_array_equal(indices.cpu().numpy(), Jo.indices)
    np.testing.assert_array_equal(data.cpu().numpy(), Jo.data)
    indptr, indices, data = f.csr(to_device(jac))
    Jg = sp.csr_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
    assert Jg.has_sorted_indices
    diff = abs(Jg - Jo).max(axis=1).toarray().ravel()
    assert np.all(diff <= 1e-12 * np.maximum(rowmax, 1e-300)), float(np.max(diff / np.maximum(rowmax, 1e-300)))


@pytest.mark.parametrize("mk", [lambda: bl_case(29, 18, seed=31),
                                lambda: cylinder_case(28, 14, seed=32),
                                lambda: cylinder_case(30, 12, seed=33),
                                lambda: bl_case(120, 61, seed=34),
                                lambda: cylinder_case(101, 40, seed=35)])
def test_fused_jacobian_csr_equals_coo_path(torch_mod, mk):
    """bca_jacobian_colored_csr (no COO) against the COO path (colored
    assembly, then the device prune + CSR, itself bit-exact against scipy on
    the oracle-pinned COO): indptr, indices and values bit for bit, including
    the retry when the value arrays are too small."""
    from broadcast_amd.flow import to_device

    case = mk()
    print(f"case {case.kind} {case.im}x{case.jm}", flush=True)
    f = _flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    cd = to_device(np.random.default_rng(36).uniform(0.5, 1.5, (case.im, case.jm)))
    jac, _, _ = f.jacobian_colored(w, coefdiag=cd)
    ref = [a.cpu().numpy() for a in f.csr(jac)]
    t = torch_mod
    n = 5 * case.im * case.jm
    f._csr_bufs = (t.empty(n + 1, dtype=t.int32, device="cuda"), t.empty(8, dtype=t.int32, device="cuda"),
                   t.empty(8, dtype=t.float64, device="cuda"))
    for _ in range(2):
        got = [a.cpu().numpy() for a in f.jacobian_csr(w, coefdiag=cd)]
        for a, b, what in zip(got, ref, ("indptr", "indices", "data")):
            np.testing.assert_array_equal(a, b, err_msg=what)
    # without the relaxation term
    jac, _, _ = f.jacobian_colored(w)
    ref = [a.cpu().numpy() for a in f.csr(jac)]
    got = [a.cpu().numpy() for a in f.jacobian_csr(w)]
    for a, b, what in zip(got, ref, ("indptr", "indices", "data")):
        np.testing.assert_array_equal(a, b, err_msg=what)


@pytest.mark.parametrize("mk", [lambda: bl_case(29, 18, seed=41),
                                lambda: cylinder_case(30, 12, seed=42),
                                lambda: cylinder_case(101, 40, seed=43)])
def test_refresh_jacobian_csr_equals_coo_path(torch_mod, mk):
    """bca_jacobian_colored_csr_refresh in a Newton-like sequence against the
    COO path at every state, bit for bit.  The first state has rho w = 0, so
    the entries that carry the spanwise velocity as a factor are exact zeros
    and pruned; the same state again leaves every kept set unchanged (data
    rewritten only); rho w perturbed makes those entries nonzero (kept sets
    grow: the CSR is rebuilt); a perturbed state; another threshold (the
    record no longer applies: full build)."""
    from broadcast_amd.flow import to_device

    t = torch_mod
    case = mk()
    print(f"case {case.kind} {case.im}x{case.jm}", flush=True)
    f = _flow(case)
    w = to_device(case.w)
    wv = w.view(5, f.nj, f.ni)
    wv[3].zero_()
    f.apply_bcs(w)
    rng = np.random.default_rng(44)
    cd = to_device(rng.uniform(0.5, 1.5, (case.im, case.jm)))

    def check(mini=2e-16):
        jac, _, _ = f.jacobian_colored(w, coefdiag=cd)
        ref = [a.cpu().numpy() for a in f.csr(jac, mini=mini)]
        got = [a.cpu().numpy() for a in f.jacobian_csr(w, coefdiag=cd, mini=mini, refresh=True)]
        for a, b, what in zip(got, ref, ("indptr", "indices", "data")):
            np.testing.assert_array_equal(a, b, err_msg=what)
        return f.csr_rebuilt, len(ref[2])

    rebuilt, nnz0 = check()
    assert rebuilt                        # no record yet: full build
    assert check() == (False, nnz0)       # same state: values only
    wi = wv[:, 3:-3, 3:-3]
    wi[3] = 1e-3 * t.tensor(rng.standard_normal(tuple(wi[3].shape)), dtype=t.float64, device="cuda")
    f.apply_bcs(w)
    rebuilt, nnz1 = check()
    assert rebuilt and nnz1 > nnz0        # rho w entries appear
    assert check() == (False, nnz1)
    wi.mul_(1.0 + 1e-3 * t.tensor(rng.standard_normal(tuple(wi.shape)), dtype=t.float64, device="cuda"))
    f.apply_bcs(w)
    check()                               # either way, the same CSR
    rebuilt, _ = check(mini=1e-6)
    assert rebuilt


def test_two_zone_seed_and_scatter_match_oracle(torch_mod, O):
    """testvector_partial + computejacobianfromjv_relaxed_withjnandcheck
    through f_misc, bit-exact, on slots prefilled with a mix of zeros and
    values (the scatter keeps or resets occupied slots)."""
    from broadcast_amd.misc import f_misc

    im, jm, gh = 17, 9, 3
    ni, nj = im + 2 * gh, jm + 2 * gh
    rng = np.random.default_rng(6)
    resd = np.asfortranarray(rng.standard_normal((ni, nj, 5)))
    cd = np.asfortranarray(rng.standard_normal((im, jm)))
    for zone in ((0, 6, 0, jm - 1), (7, im - 1, 0, jm - 1), (3, 12, 2, 6)):
        a, b = np.zeros((ni, nj, 5), order="F"), np.zeros((ni, nj, 5), order="F")
        f_misc.testvector_partial(a, 2, 4, 1, gh, im, jm, *zone)
        O.testvector_partial(b, 2, 4, 1, gh, im, jm, *zone)
        np.testing.assert_array_equal(a, b)
    nb = im * jm * 49 * 25 * 2
    pre = np.where(rng.random(nb) < 0.3, rng.standard_normal(nb), 0.0)
    g = [pre.copy(), np.full(nb, 7, np.int32), np.full(nb, 9, np.int32)]
    o = [pre.copy(), np.full(nb, 7, np.int32), np.full(nb, 9, np.int32)]
    for n in (0, 1):
        for m, l, k in [(0, 0, 0), (2, 5, 3), (4, 6, 6), (1, 3, 0), (3, 2, 5)]:
            f_misc.computejacobianfromjv_relaxed_withjnandcheck(*g, resd, m, l, k, gh, cd, 2e-16, n)
            O.computejacobianfromjv_relaxed_withjnandcheck(*o, resd, m, l, k, gh, im, jm, cd, 2e-16, n)
    for a, b in zip(g, o):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mk", [lambda: cylinder_case(30, 12, seed=61), lambda: bl_case(29, 18, seed=62)])
def test_jacobian_graph_replays_match_eager(torch_mod, mk):
    """The Jacobian stage captured in a HIP graph: every replay, with eager
    work and a changed state and relaxation in between, gives the eager
    stage's residual and COO values bit for bit."""
    from broadcast_amd.flow import to_device
    from broadcast_amd.graph import JacobianGraph

    t = torch_mod
    case = mk()
    f = _flow(case)
    w = to_device(case.w)
    cd = to_device(np.random.default_rng(63).uniform(0.5, 1.5, (case.im, case.jm)))
    G = JacobianGraph(f, w, coefdiag=cd)
    rng = np.random.default_rng(64)
    for rep in range(3):
        # change the captured inputs in place, then replay
        w.copy_(to_device(case.w) * (1.0 + 1e-3 * rep))
        cd.copy_(to_device(rng.uniform(0.5, 1.5, (case.im, case.jm))))
        jac_g = G.replay()[0].clone()
        res_g = G.res.clone()
        # eager reference on copies of the same inputs
        we = w.clone()
        f.apply_bcs(we)
        res_e = f.residual(we)
        jac_e, _, _ = f.jacobian_colored(we, coefdiag=cd)
        t.cuda.synchronize()
        assert t.equal(jac_g, jac_e), f"replay {rep}: COO values differ"
        assert t.equal(res_g.view(5, f.nj, f.ni)[:, 3:-3, 3:-3], res_e.view(5, f.nj, f.ni)[:, 3:-3, 3:-3])
        assert t.equal(w, we), f"replay {rep}: ghost states differ"


@pytest.mark.parametrize("mk", [lambda: cylinder_case(44, 20, seed=71), lambda: bl_case(40, 24, seed=72)])
def test_autotune_keeps_results(torch_mod, mk):
    """Launch shapes chosen by bca_autotune change no value: residual and
    colored COO before and after tuning are identical."""
    from broadcast_amd.flow import to_device

    t = torch_mod
    case = mk()
    f = _flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    cd = to_device(np.random.default_rng(73).uniform(0.5, 1.5, (case.im, case.jm)))
    r0 = f.residual(w).clone()
    j0 = f.jacobian_colored(w, coefdiag=cd)[0].clone()
    rep = f.autotune(w)
    print(rep, flush=True)
    assert "k_face" in rep and "k_assemble_colored" in rep
    r1 = f.residual(w)
    j1 = f.jacobian_colored(w, coefdiag=cd)[0]
    t.cuda.synchronize()
    assert t.equal(r0, r1) and t.equal(j0, j1)


@pytest.mark.parametrize("name", CASES)
def test_norml2_and_set_bndbl_match_reference(torch_mod, name):
    """compute_norml2 (srcfv/norm.F90:2-32) and set_bndbl_2d (set_bnd.f90:2-24)
    through the f2py call forms (BROADCAST_npz.py:773; srcfv/f_norm.py:24), on
    numpy arrays and on device tensors, against the reference's own outputs
    in the fixtures: norms to 1e-13, the copies bit-exact."""
    from broadcast_amd import f_init
    from broadcast_amd.flow import to_device
    from broadcast_amd.srcfv import f_norm

    fx = load(name)
    im, jm, gh = int(fx["im"]), int(fx["jm"]), int(fx["gh"])
    for res in (np.array(fx["res"], order="F"), to_device(fx["res"])):
        norm, nmoy = f_norm.compute_norml2(res, im, jm, gh)
        np.testing.assert_allclose(norm, fx["norm"], rtol=1e-13)
        np.testing.assert_allclose(nmoy, fx["nmoy"], rtol=1e-13)
    w = np.array(fx["w_bc"], order="F")
    field = np.full((jm, gh, 5), np.nan, order="F")
    wbd = np.full((im + gh, 5), np.nan, order="F")
    f_init.set_bndbl_2d(w, field, wbd, im)
    np.testing.assert_array_equal(field, fx["sbd_field"])
    np.testing.assert_array_equal(wbd, fx["sbd_wbd"])
    np.testing.assert_array_equal(w, fx["w_bc"])


def test_full_size_cylinder_csr_applies_the_jacobian(torch_mod):
    """C1 at its configured size: the cylinder O-mesh of 512 x 384 cells (two
    i-zones, 490 seeds).  The pruned CSR of the colored Jacobian applied to 3
    seeded vectors u equals the pinned dense path -(BC o R)'(w) u (tangent
    BCs + JVP), 1e-12 normwise per variable.  With the relaxation term the
    CSR differs from it on the diagonal only, by mult * coefdiag: the
    two-zone loop adds coefdiag once per zone whose window holds the
    diagonal (the multiplicity pinned against the reference build's
    scatter in the CPU suite), and the fused path gives the same CSR."""
    t = torch_mod
    from broadcast_amd.flow import Flow, to_device
    from golden_util import two_zone_diag_multiplicity

    case = cylinder_case(512, 384, seed=1000)
    im, jm = case.im, case.jm
    print(f"C1 full size {im}x{jm}", flush=True)
    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    rng = np.random.default_rng(1001)
    cd_np = rng.uniform(0.5, 1.5, (im, jm))
    cd = t.tensor(cd_np.ravel(order="F"), dtype=t.float64, device="cuda")
    n = 5 * im * jm

    def coo_csr(coefdiag):
        jac, ia, ja = f.jacobian_colored(w, coefdiag=coefdiag)
        out = f.csr(jac)
        del jac, ia, ja
        return out

    indptr, indices, data = coo_csr(None)
    assert int(indptr[-1]) == data.numel() and data.numel() > 100 * n
    rows = t.repeat_interleave(t.arange(n, device="cuda"), (indptr[1:] - indptr[:-1]).long())
    cols = indices.long()
    for s in range(3):
        u = rng.standard_normal((im, jm, 5))
        ut = t.tensor(u.ravel(), dtype=t.float64, device="cuda")     # unknown e + 5 (j + jm i)
        y = t.zeros(n, dtype=t.float64, device="cuda").index_add_(0, rows, data * ut[cols])
        wd = np.zeros(f.shape_w, order="F")
        wd[3:-3, 3:-3, :] = u
        _, resd = f.jvp(w.clone(), to_device(wd))
        ref = -_host(f, resd, f.shape_w)[3:-3, 3:-3, :]
        ok, err = close(y.cpu().numpy().reshape(im, jm, 5), ref, 1e-12)
        assert ok, f"direction {s}: CSR u vs -(BC o R)' u, rel err per variable {err}"
    # the relaxed matrix: same pattern, diagonal + mult * coefdiag
    ip2, ix2, d2 = coo_csr(cd)
    assert t.equal(ip2, indptr) and t.equal(ix2, indices)
    diag = rows == cols
    assert int(diag.sum()) == n
    off = ~diag
    assert t.equal(d2[off], data[off])
    mult = two_zone_diag_multiplicity(im)                             # per i
    want = (mult[:, None] * cd_np)[:, :, None].repeat(5, axis=2).ravel()  # per row e + 5 (j + jm i)
    got = (d2[diag] - data[diag]).cpu().numpy()
    scale = np.abs(d2[diag].cpu().numpy()) + np.abs(data[diag].cpu().numpy())
    assert np.all(np.abs(got - want) <= 1e-14 * scale + 1e-15), float(np.max(np.abs(got - want) / scale))
    # the fused path without the COO gives this CSR bit for bit
    fused = f.jacobian_csr(w, coefdiag=cd)
    for a, b, what in zip(fused, (ip2, ix2, d2), ("indptr", "indices", "data")):
        assert t.equal(a, b), what


@pytest.mark.parametrize("kind", ["extrap", "wall"])
def test_tangent_bc_frame_and_full_array_packing(torch_mod, O, kind):
    """A tangent BC whose box lies on the block edge packs and unpacks the
    boundary frame only; one whose box lies inside the block (column im-6 or
    row 6 of a 24 x 16 block) takes the whole-array pack.  Both give the
    oracle's tangent BC on the same arrays, and leave every cell the BC does
    not write untouched."""
    from broadcast_amd.srcfv import f_lin

    case = bl_case(24, 16, name="bl24x16")
    im, jm, gh = case.im, case.jm, case.gh
    rng = np.random.default_rng(11)
    shape = (im + 2 * gh, jm + 2 * gh, 5)
    w0 = np.asfortranarray(1.0 + 0.1 * rng.random(shape))
    w0[..., 4] += 2.5  # positive pressure for the wall's power law
    wd0 = np.asfortranarray(rng.standard_normal(shape))
    if kind == "extrap":
        loc, boxes = "Ihi", [case.interf["Ihi"], np.array([[im - 6, 1], [im - 6, jm]], dtype=float, order="F")]
    else:
        loc, boxes = "Jlo", [case.interf["Jlo"], np.array([[1, 6], [im, 6]], dtype=float, order="F")]
    for box in boxes:
        w, wd = w0.copy(order="F"), wd0.copy(order="F")
        if kind == "extrap":
            f_lin.bc_extrapolate_o2_2d_d(w, wd, loc, box, im, jm, gh)
        else:
            f_lin.bc_wall_viscous_adia_2d_d(w, wd, loc, case.phys.gam, box, gh, im, jm)
        wr, wdr = w0.copy(order="F"), wd0.copy(order="F")
        if kind == "extrap":
            O.bc_extrapolate_o2_2d(wr, loc, box, im, jm, gh, wd=wdr)
        else:
            O.bc_wall_viscous_adia_2d(wr, loc, case.phys.gam, box, gh, im, jm, wd=wdr)
        assert not np.array_equal(wr, w0), "the BC wrote nothing"
        ok, err = close(w, wr, 1e-14)
        assert ok, (box.tolist(), err)
        ok, err = close(wd, wdr, 1e-13)
        assert ok, (box.tolist(), err)
        untouched = (wr == w0) & (wdr == wd0)
        assert np.array_equal(w[untouched], w0[untouched]), box.tolist()
        assert np.array_equal(wd[untouched], wd0[untouched]), box.tolist()
