"""NPZ tree (broadcast_amd.npz) against the format the reference writes and
reads (BROADCAST_npz.py:247-262, 311-333, 407-467): keys, dtypes and shapes,
the reader's string handling and end-of-run preference, and a restart: the
case rebuilt from a tree gives the same residual (oracle) as the original."""
import numpy as np

import oracle
from oracle import driver as D
from broadcast_amd import npz
from broadcast_amd.cases import bl_case

# every key writeNPZ stores (BROADCAST_npz.py:261)
WRITE_KEYS = {"x0", "y0", "xc", "yc", "vol", "volf", "nx", "ny", "im", "jm", "gh", "FlowSolutionInit",
              "ResidualInit", "interf1", "interf2", "interf3", "interf4", "BC1", "BC2", "BC3", "BC4", "Wbd",
              "Field", "SchemeType", "Scheme", "k2", "k4", "Mach", "Prandtl", "Rgaz", "Cp", "Cv", "Gamma",
              "TemperatureSutherland", "TemperatureRefSutherland", "ViscosityRefSutherland", "Pref", "CFL",
              "Freqres", "Freqsort"}
FILL_KEYS = {"ResidualEndOfRun", "FlowSolutionEndOfRun", "IA", "JA", "Aij"}
FILL3D_KEYS = {"IAdz", "JAdz", "Aijdz", "IAdz2", "JAdz2", "Aijdz2"}


def _write(tmp_path, case, O):
    g = D.geometry(O, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(O, case, g, w)
    res = D.residual(O, case, g, w)
    ph = case.phys
    cp, cv, prandtl, gam, rgaz, cs, muref, tref, _, k2, k4 = ph.scheme_args()
    it = case.interf
    fn = str(tmp_path / "tree")
    npz.writeNPZ(fn, case.im, case.jm, case.gh, w, g["x0"], g["y0"], g["vol"], g["volf"], g["nx"], g["ny"],
                 g["xc"], g["yc"], case.field, case.wbd, res, "dnc", k2, k4, it["Ilo"], it["Ihi"], it["Jlo"],
                 it["Jhi"], npz.flow_lf(case), cp, cv, gam, cs, tref, muref, rgaz, ph.mach, prandtl, 1e10, 1, 10,
                 pinf=1.0)
    return fn, g, w, res


def test_tree_keys_dtypes_and_reader(tmp_path):
    O = oracle.Oracle()
    case = bl_case(20, 12, seed=7)
    fn, g, w, res = _write(tmp_path, case, O)
    d = np.load(fn + ".npz", allow_pickle=False)
    assert set(d.files) == WRITE_KEYS
    ni, nj = case.ni, case.nj
    shapes = {"x0": (ni + 1, nj + 1), "y0": (ni + 1, nj + 1), "xc": (ni, nj), "yc": (ni, nj), "vol": (ni, nj),
              "volf": (ni, nj, 2), "nx": (ni + 1, nj + 1, 2), "ny": (ni + 1, nj + 1, 2),
              "FlowSolutionInit": (ni, nj, 5), "ResidualInit": (ni, nj, 5), "interf1": (2, 2),
              "Field": (case.jm, case.gh, 5), "Wbd": case.wbd.shape}
    for k, s in shapes.items():
        assert d[k].shape == s and d[k].dtype == np.float64, k
    for k in ("BC1", "BC2", "BC3", "BC4", "Scheme", "SchemeType"):
        assert d[k].dtype.kind == "U" and d[k].ndim == 0, k
    assert str(d["BC1"]) == "f_bnd.bc_supandsubinlet_2d" and str(d["BC4"]) == "f_bnd.bc_no_reflexion_2d"
    assert str(d["Scheme"]) == "f_sch.flux_num_dnc5_2d"
    np.testing.assert_array_equal(d["FlowSolutionInit"], w)
    t = npz.readNPZtree(fn)
    assert len(t) == 36
    assert t[15] == "dnc"
    assert t[22] == ["bc_supandsubinlet_2d", "bc_extrapolate_o2_2d", "bc_no_reflexion_2d",
                     "bc_wall_viscous_adia_2d", "flux_num_dnc5_2d", "f_bnd", "f_sch"]
    np.testing.assert_array_equal(t[3], w)          # initial state while no end-of-run state
    # fill: end-of-run state preferred by the reader, Jacobian lists stored as given
    rng = np.random.default_rng(3)
    IA = rng.integers(0, 5 * case.im * case.jm, 50).astype(np.int32)
    JA = rng.integers(0, 5 * case.im * case.jm, 50).astype(np.int32)
    A = rng.standard_normal(50)
    w2 = w + 1.0
    npz.fillNPZ(fn, w2, res, IA, JA, A, case.gh)
    npz.fillNPZ_3D(fn, IA, JA, A, IA[:7], JA[:7], A[:7])
    d = np.load(fn + ".npz", allow_pickle=False)
    assert set(d.files) == WRITE_KEYS | FILL_KEYS | FILL3D_KEYS
    assert d["IA"].dtype == np.int32 and d["JA"].dtype == np.int32 and d["Aij"].dtype == np.float64
    np.testing.assert_array_equal(d["Aijdz2"], A[:7])
    t = npz.readNPZtree(fn)
    np.testing.assert_array_equal(t[3], w2)


def test_restart_from_tree_gives_the_same_residual(tmp_path):
    O = oracle.Oracle()
    case = bl_case(20, 12, seed=8)
    fn, g, w, res = _write(tmp_path, case, O)
    c2 = npz.case_from_tree(fn)
    g2 = D.geometry(O, c2)
    for k in ("nx", "ny", "vol", "volf"):
        np.testing.assert_array_equal(g2[k], g[k], err_msg=k)
    w2 = np.array(c2.w, order="F", copy=True)
    D.apply_bcs(O, c2, g2, w2)
    np.testing.assert_array_equal(w2, w)
    np.testing.assert_array_equal(D.residual(O, c2, g2, w2), res)


def test_remove_zero_jac_keeps_slot_order():
    jac = np.array([0.0, 1.0, -1e-17, 3e-16, -2.0, 2e-16])
    ia = np.arange(6, dtype=np.int32)
    ja = ia[::-1].copy()
    IA, JA, J = npz.remove_zero_jac(ia, ja, jac)
    np.testing.assert_array_equal(IA, [1, 3, 4])
    np.testing.assert_array_equal(JA, [4, 2, 1])
    np.testing.assert_array_equal(J, [1.0, 3e-16, -2.0])
