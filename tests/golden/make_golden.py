"""Writes the golden vectors of the FP64 residual / tangent / adjoint path.

Every expected output here is produced by the reference Fortran itself,
compiled from /root/reference by oracle/Makefile.ref into
oracle/_ref/libbroadcast_ref.so and called through its Fortran ABI.  Inputs
are the seeded synthetic cases of broadcast_amd.cases.  Run from the
repository root:

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

`--dz-only` adds (or refreshes) the spanwise operators dz, dz2 of the case
fixtures from their stored inputs without regenerating anything else;
`--bcb-only` likewise adds the boundary-condition adjoints (bc_*_b),
`--sbd-only` the outputs of set_bndbl_2d and `--nowall-only` the residual and
tangent of the variant without wall rows (flux_num_dnc5_nowall_2d[_d]);
`--dnc3-only` writes the 3rd-order scheme's fixtures (dnc3_*.npz, gh = 2),
`--dnc79-only` those of orders 7 and 9 (dnc7_*.npz, gh = 4; dnc9_*.npz,
gh = 5).

Fixtures (tests/golden/*.npz) hold inputs and expected outputs only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from broadcast_amd.cases import bl_case, cylinder_case  # noqa: E402
import oracle  # noqa: E402
from oracle import driver as D  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def case_fixture(R, case, seed):
    g = D.geometry(R, case)
    w_in = np.array(case.w, order="F", copy=True)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(R, case, g, w)
    res = D.residual(R, case, g, w)
    norm, ninf = R.compute_norml2inf(res, case.im, case.jm, case.gh)
    norm2, nmoy = R.compute_norml2(res, case.im, case.jm, case.gh)
    rng = np.random.default_rng(seed)
    gh = case.gh
    wd = np.zeros((2,) + w.shape)
    resd = np.zeros((2,) + w.shape)
    wd_bc = np.zeros((2,) + w.shape)
    for s in range(2):
        d = np.zeros_like(w, order="F")
        d[gh:-gh, gh:-gh, :] = rng.standard_normal((case.im, case.jm, 5))
        wd[s] = d
        wcopy = np.array(w, order="F", copy=True)
        dcopy = np.array(d, order="F", copy=True)
        D.apply_bcs(R, case, g, wcopy, wd=dcopy)
        wd_bc[s] = dcopy
        _, rd = D.jvp(R, case, g, w, d)
        resd[s] = rd
    # adjoint of the flux routine alone (no BC adjoint): wb = (dR/dw)^T resb
    resb = np.zeros_like(w, order="F")
    resb[gh:-gh, gh:-gh, :] = rng.standard_normal((case.im, case.jm, 5))
    resb_in = resb.copy(order="F")
    wb = np.zeros_like(w, order="F")
    r2 = np.zeros_like(w, order="F")
    R.flux_num_dnc5_2d_b(r2, resb, w, wb, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"],
                         g["vol"], g["volf"], gh, *case.phys.scheme_args(), case.im, case.jm)
    ph = case.phys
    return dict(
        kind=np.array(case.kind), im=case.im, jm=case.jm, gh=gh,
        phys=np.array(ph.scheme_args()), mach=ph.mach, reynolds=ph.reynolds,
        x0_in=case.x0, y0_in=case.y0, w_in=w_in,
        field=case.field if case.field is not None else np.zeros((0,)),
        wbd=case.wbd,
        **{f"interf_{k}": v for k, v in case.interf.items()},
        **{f"geo_{k}": v for k, v in g.items()},
        w_bc=w, res=res, norm=norm, ninf=ninf, nmoy=nmoy,
        wd=wd, wd_bc=wd_bc, resd=resd, resb=resb_in, wb=wb,
    )


def dz_vectors(R, fx):
    """coeffs_5p_dz / coeffs_5p_dz2 (srcfv/dz) on the stored state after BCs
    and the first tangent direction after the linearised BCs."""
    g = {k: np.asfortranarray(fx[f"geo_{k}"]) for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    w = np.asfortranarray(fx["w_bc"])
    wd = np.asfortranarray(fx["wd_bc"][0])
    args = tuple(float(v) for v in fx["phys"][:9])
    gh, im, jm = int(fx["gh"]), int(fx["im"]), int(fx["jm"])
    dz = np.zeros_like(w, order="F")
    dz2 = np.zeros_like(w, order="F")
    for fn, out in ((R.coeffs_5p_dz, dz), (R.coeffs_5p_dz2, dz2)):
        fn(out, w, wd, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], gh, *args,
           im, jm)
    return dict(dz=dz, dz2=dz2)


def bcb_vectors(R, fx, seed=21):
    """Each BC adjoint of the case's sequence (srcfv/adjoint/bc_*_b.f90) on
    the stored state after BCs and one seeded full-array adjoint bcb_wb:
    bcb_<loc> = wb after that routine alone."""
    g = {k: np.asfortranarray(fx[f"geo_{k}"]) for k in ("nx", "ny")}
    w = np.asfortranarray(fx["w_bc"])
    gh, im, jm = int(fx["gh"]), int(fx["im"]), int(fx["jm"])
    gam = float(fx["phys"][3])
    it = {k: fx[f"interf_{k}"] for k in ("Ilo", "Ihi", "Jlo", "Jhi")}
    rng = np.random.default_rng(seed)
    wb0 = np.asfortranarray(rng.standard_normal(w.shape))
    calls = {
        "Jhi": lambda a, b: R.bc_no_reflexion_2d_b(a, b, np.asfortranarray(fx["wbd"]), "Jhi", it["Jhi"], g["nx"],
                                                   g["ny"], gam, gh, im, jm),
        "Jlo": lambda a, b: R.bc_wall_viscous_adia_2d_b(a, b, "Jlo", gam, it["Jlo"], gh, im, jm),
    }
    if str(fx["kind"]) == "bl":
        calls["Ilo"] = lambda a, b: R.bc_supandsubinlet_2d_b(a, b, "Ilo", it["Ilo"], np.asfortranarray(fx["field"]),
                                                             g["nx"], g["ny"], gam, im, jm, gh)
        calls["Ihi"] = lambda a, b: R.bc_extrapolate_o2_2d_b(a, b, "Ihi", it["Ihi"], im, jm, gh)
    out = {"bcb_wb": wb0}
    for loc, fn in calls.items():
        a = np.array(w, order="F", copy=True)
        b = np.array(wb0, order="F", copy=True)
        fn(a, b)
        assert np.array_equal(a, w), "the reference restores w"
        out[f"bcb_{loc}"] = b
    return out


def sbd_vectors(R, fx):
    """set_bndbl_2d (set_bnd.f90:2-24) on the stored state after BCs: the
    inflow profile field(jm, gh, 5) and the top-row values wbd(im+gh, 5),
    starting from seeded non-zero buffers (the routine overwrites them)."""
    w = np.asfortranarray(fx["w_bc"])
    gh, im, jm = int(fx["gh"]), int(fx["im"]), int(fx["jm"])
    rng = np.random.default_rng(31)
    field = np.asfortranarray(rng.standard_normal((jm, gh, 5)))
    wbd = np.asfortranarray(rng.standard_normal((im + gh, 5)))
    R.set_bndbl_2d(w, field, wbd, im, jm, gh)
    return dict(sbd_field=field, sbd_wbd=wbd)


def nowall_vectors(R, fx):
    """The variant without wall rows (srcfv/rhs/flux_num_dnc5_nowall.F90 and
    its tangent srcfv/tangent/flux_num_dnc5_nowall_d.f90) on the stored state
    after BCs and the first tangent direction after the linearised BCs:
    res_nowall, resd_nowall."""
    g = {k: np.asfortranarray(fx[f"geo_{k}"]) for k in ("x0", "y0", "nx", "ny", "xc", "yc", "vol", "volf")}
    w = np.asfortranarray(fx["w_bc"])
    wd = np.asfortranarray(fx["wd_bc"][0])
    args = tuple(float(v) for v in fx["phys"])
    gh, im, jm = int(fx["gh"]), int(fx["im"]), int(fx["jm"])
    res = np.zeros_like(w, order="F")
    R.flux_num_dnc5_nowall_2d(res, np.array(w, order="F", copy=True), g["x0"], g["y0"], g["nx"], g["ny"], g["xc"],
                              g["yc"], g["vol"], g["volf"], gh, *args, im, jm)
    r2 = np.zeros_like(w, order="F")
    resd = np.zeros_like(w, order="F")
    R.flux_num_dnc5_nowall_2d_d(r2, resd, np.array(w, order="F", copy=True), np.array(wd, order="F", copy=True),
                                g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], gh,
                                *args, im, jm)
    # the generated tangent assigns only residud: residu is left as given
    # (flux_num_dnc5_2d_d behaves the same way)
    assert not r2.any()
    return dict(res_nowall=res, resd_nowall=resd)


def order_fixture(R, case, seed, order):
    """The scheme of another order on a case of gh = (order + 1) / 2
    (BROADCAST_npz.py:501-502): flux_num_dnc{order}_2d and
    flux_num_dnc{order}_nowall_2d (srcfv/rhs) on the state after the case's
    BCs, and their tangents (srcfv/tangent/flux_num_dnc{order}*_d.f90) along
    one seeded direction after the linearised BCs: res{order}[_nowall],
    resd{order}[_nowall]."""
    assert case.gh == (order + 1) // 2
    g = D.geometry(R, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(R, case, g, w)
    rng = np.random.default_rng(seed)
    gh, im, jm = case.gh, case.im, case.jm
    d = np.zeros_like(w, order="F")
    d[gh:-gh, gh:-gh, :] = rng.standard_normal((im, jm, 5))
    wd_in = d.copy(order="F")
    wcopy = np.array(w, order="F", copy=True)
    D.apply_bcs(R, case, g, wcopy, wd=d)
    args = case.phys.scheme_args()
    out = {}
    for tag, nowall in (("", False), ("_nowall", True)):
        fn, fnd = R.scheme(order, nowall, False), R.scheme(order, nowall, True)
        res = np.zeros_like(w, order="F")
        fn(res, np.array(w, order="F", copy=True), g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"],
           g["vol"], g["volf"], gh, *args, im, jm)
        r2 = np.zeros_like(w, order="F")
        resd = np.zeros_like(w, order="F")
        fnd(r2, resd, np.array(w, order="F", copy=True), np.array(d, order="F", copy=True), g["x0"], g["y0"],
            g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], gh, *args, im, jm)
        assert not r2.any()  # the generated tangent assigns residud only
        out[f"res{order}{tag}"] = res
        out[f"resd{order}{tag}"] = resd
    return dict(kind=np.array(case.kind), im=im, jm=jm, gh=gh, phys=np.array(args),
                x0_in=case.x0, y0_in=case.y0, w_in=case.w,
                field=case.field if case.field is not None else np.zeros((0,)),
                wbd=case.wbd, **{f"interf_{k}": v for k, v in case.interf.items()},
                **{f"geo_{k}": v for k, v in g.items()},
                w_bc=w, wd=wd_in, wd_bc=d, **out)


def jac_fixture(R, case):
    g = D.geometry(R, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(R, case, g, w)
    rng = np.random.default_rng(77)
    coefdiag = np.asfortranarray(rng.uniform(0.5, 1.5, (case.im, case.jm)))
    J = D.colored_jacobian(R, case, g, w, coefdiag=coefdiag)
    J.sort_indices()
    return dict(kind=np.array(case.kind), im=case.im, jm=case.jm, gh=case.gh,
                phys=np.array(case.phys.scheme_args()), x0_in=case.x0, y0_in=case.y0,
                w_in=case.w, field=case.field if case.field is not None else np.zeros((0,)),
                wbd=case.wbd, **{f"interf_{k}": v for k, v in case.interf.items()},
                w_bc=w, coefdiag=coefdiag, indptr=J.indptr.astype(np.int32),
                indices=J.indices.astype(np.int32), data=J.data)


CASES = ("bl24x16", "bl13x9", "cyl28x12")


def DNC3_CASES(R):
    return {"dnc3_bl20x12": order_fixture(R, bl_case(20, 12, seed=6, gh=2), 16, 3),
            "dnc3_cyl21x10": order_fixture(R, cylinder_case(21, 10, seed=7, gh=2), 17, 3)}


def DNC79_CASES(R):
    return {"dnc7_bl22x14": order_fixture(R, bl_case(22, 14, seed=8, gh=4), 18, 7),
            "dnc7_cyl24x12": order_fixture(R, cylinder_case(24, 12, seed=9, gh=4), 19, 7),
            "dnc9_bl23x16": order_fixture(R, bl_case(23, 16, seed=10, gh=5), 20, 9),
            "dnc9_cyl26x13": order_fixture(R, cylinder_case(26, 13, seed=11, gh=5), 21, 9)}


def main():
    R = oracle.Reference()
    only = {"--dz-only": dz_vectors, "--bcb-only": bcb_vectors, "--sbd-only": sbd_vectors,
            "--nowall-only": nowall_vectors}
    extra = [fn for flag, fn in only.items() if flag in sys.argv]
    if "--dnc3-only" in sys.argv or "--dnc79-only" in sys.argv:
        cases = DNC3_CASES(R) if "--dnc3-only" in sys.argv else DNC79_CASES(R)
        for name, fx in cases.items():
            path = os.path.join(OUT, f"{name}.npz")
            np.savez_compressed(path, **fx)
            print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")
        return
    if extra:
        for name in CASES:
            path = os.path.join(OUT, f"{name}.npz")
            fx = dict(np.load(path, allow_pickle=False))
            for fn in extra:
                fx.update(fn(R, fx))
            np.savez_compressed(path, **fx)
            print(f"updated {path} ({os.path.getsize(path) / 1024:.0f} KiB)")
        return
    fixtures = {
        "bl24x16": case_fixture(R, bl_case(24, 16, seed=1), 11),
        "bl13x9": case_fixture(R, bl_case(13, 9, seed=3, reynolds=3e5), 13),
        "cyl28x12": case_fixture(R, cylinder_case(28, 12, seed=2), 12),
        "jac_bl14x10": jac_fixture(R, bl_case(14, 10, seed=4)),
        "jac_cyl14x8": jac_fixture(R, cylinder_case(14, 8, seed=5)),
    }
    for name in CASES:
        fixtures[name].update(dz_vectors(R, fixtures[name]))
        fixtures[name].update(bcb_vectors(R, fixtures[name]))
        fixtures[name].update(sbd_vectors(R, fixtures[name]))
        fixtures[name].update(nowall_vectors(R, fixtures[name]))
    fixtures.update(DNC3_CASES(R))
    fixtures.update(DNC79_CASES(R))
    for name, fx in fixtures.items():
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **fx)
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
