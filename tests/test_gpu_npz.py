"""NPZ tree lists from device results: the Jacobian lists of the reference's
post-Newton step (remove_zero_jac in slot order, Aij = Jac / vol of the row
cell, BROADCAST_npz.py:1137, 1204-1215) and the Dz / Dzz lists
(:1248-1253), bit-exact against the same steps done in numpy on the device
COO (itself checked against the oracle in test_gpu_parity / test_gpu_dz),
written with fillNPZ / fillNPZ_3D and read back by readNPZtree."""
import numpy as np
import pytest

from broadcast_amd import npz
from broadcast_amd.cases import bl_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


def test_device_lists_round_trip(torch_mod, tmp_path):
    t = torch_mod
    from broadcast_amd.flow import Flow, to_device, to_host

    case = bl_case(24, 16, seed=9)
    im, jm, gh = case.im, case.jm, case.gh
    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    res = f.residual(w)
    cd = t.tensor(np.random.default_rng(4).uniform(0.5, 1.5, im * jm), dtype=t.float64, device="cuda")
    jac, ia, ja = f.jacobian_colored(w, coefdiag=cd)
    IA, JA, Jv = npz.jacobian_lists(f, jac, ia, ja)
    jh, iah, jah = jac.cpu().numpy(), ia.cpu().numpy(), ja.cpu().numpy()
    keep = np.abs(jh) > 2e-16
    vol = to_host(f.vol, (f.ni, f.nj))
    IAh, JAh = iah[keep], jah[keep]
    Jvh = jh[keep] / vol[IAh // (jm * 5) + gh, (IAh % (jm * 5)) // 5 + gh]
    np.testing.assert_array_equal(IA.cpu().numpy(), IAh)
    np.testing.assert_array_equal(JA.cpu().numpy(), JAh)
    np.testing.assert_array_equal(Jv.cpu().numpy(), Jvh)
    jdz, jdz2, iadz, jadz = f.dz_colored(w)
    lists = npz.dz_lists(jdz, jdz2, iadz, jadz)
    for (I, J, A), vals in ((lists[:3], jdz), (lists[3:], jdz2)):
        v = vals.cpu().numpy()
        k = np.abs(v) > 2e-16
        np.testing.assert_array_equal(I.cpu().numpy(), iadz.cpu().numpy()[k])
        np.testing.assert_array_equal(J.cpu().numpy(), jadz.cpu().numpy()[k])
        np.testing.assert_array_equal(A.cpu().numpy(), v[k])
    fn = str(tmp_path / "tree")
    g = {k: getattr(f, k) for k in ("nx", "ny", "xc", "yc", "vol", "volf")}
    cp, cv, prandtl, gam, rgaz, cs, muref, tref, _, k2, k4 = case.phys.scheme_args()
    it = case.interf
    npz.writeNPZ(fn, im, jm, gh, w, case.x0, case.y0, g["vol"], g["volf"], g["nx"], g["ny"], g["xc"], g["yc"],
                 case.field, case.wbd, res, "dnc", k2, k4, it["Ilo"], it["Ihi"], it["Jlo"], it["Jhi"],
                 npz.flow_lf(case), cp, cv, gam, cs, tref, muref, rgaz, case.phys.mach, prandtl, 1e10, 1, 10)
    npz.fillNPZ(fn, w, res, IA, JA, Jv, gh)
    npz.fillNPZ_3D(fn, *lists)
    d = np.load(fn + ".npz", allow_pickle=False)
    np.testing.assert_array_equal(d["IA"], IAh)
    np.testing.assert_array_equal(d["Aij"], Jvh)
    np.testing.assert_array_equal(d["FlowSolutionEndOfRun"], to_host(w, f.shape_w))
    assert d["IAdz"].dtype == np.int32 and d["Aijdz2"].dtype == np.float64
    tree = npz.readNPZtree(fn)
    np.testing.assert_array_equal(tree[3], to_host(w, f.shape_w))
