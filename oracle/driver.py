"""ORACLE -- test infrastructure only.

Driver-level sequences of the reference, expressed over either checker
(``oracle.Oracle`` or ``oracle.Reference``): geometry set-up, boundary-condition
order, residual, tangent JVP and the colored-seed Jacobian assembly.

Follows BROADCAST_npz.py:702 (geometry), :1018-1031 (BC order + residual),
:1068-1127 (245-seed Jacobian loop) for the flat-plate block, and
cylinder.py:496-527 + handleBC.py:129-242 (join, BC order) for the O-mesh.
"""
from __future__ import annotations

import numpy as np

from broadcast_amd.cases import Case, join_boxes


def geometry(lib, case: Case):
    im, jm, gh = case.im, case.jm, case.gh
    ni, nj = case.ni, case.nj
    g = dict(
        x0=np.array(case.x0, order="F", copy=True),
        y0=np.array(case.y0, order="F", copy=True),
        nx=np.zeros((ni + 1, nj + 1, 2), order="F"),
        ny=np.zeros((ni + 1, nj + 1, 2), order="F"),
        xc=np.zeros((ni, nj), order="F"),
        yc=np.zeros((ni, nj), order="F"),
        vol=np.zeros((ni, nj), order="F"),
        volf=np.zeros((ni, nj, 2), order="F"),
    )
    lib.computegeom_2d(g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"],
                       g["volf"], im, jm, gh)
    if case.kind == "cyl":
        # cylinder.py:499-527: periodic ghost metrics replace the linear
        # extrapolation of computegeom_2d.
        (prr1, prd2, tr), (prr2, prd1, _) = join_boxes(im, jm, gh)
        for key in ("xc", "yc", "vol"):
            a = g[key][:, :, None].copy(order="F")
            lib.jn_match_2d(a, prr1, gh, gh, gh, gh, im, jm, a, prd2, gh, gh, gh, gh, im, jm, tr)
            lib.jn_match_2d(a, prr2, gh, gh, gh, gh, im, jm, a, prd1, gh, gh, gh, gh, im, jm, tr)
            g[key][:, :] = a[:, :, 0]
        jmin, jmax = 1 - gh, jm + gh

        def box(a, b, c, d):
            return np.asfortranarray(np.array([[a, b], [c, d]], dtype=np.int32))

        prr1g = box(im + 2, jmin, im + 1 + gh, jmax + 1)
        prd1g = box(im + 1 - gh, jmin, im, jmax + 1)
        prr2g = box(1 - gh, jmin, 0, jmax + 1)
        prd2g = box(2, jmin, 1 + gh, jmax + 1)
        for key in ("nx", "ny"):
            a = g[key]
            lib.jn_match_2d(a, prr1g, gh, gh, gh, gh, im + 1, jm + 1, a, prd2g, gh, gh, gh, gh,
                            im + 1, jm + 1, tr)
            lib.jn_match_2d(a, prr2g, gh, gh, gh, gh, im + 1, jm + 1, a, prd1g, gh, gh, gh, gh,
                            im + 1, jm + 1, tr)
        vol, volf = g["vol"], g["volf"]
        for j in range(gh, jm + 1 + gh):
            for i in range(gh, im + 1 + gh):
                volf[i, j, 0] = 2.0 / (vol[i, j] + vol[i - 1, j])
                volf[i, j, 1] = 2.0 / (vol[i, j] + vol[i, j - 1])
    return g


def apply_bcs(lib, case: Case, g, w, wd=None):
    """Ghost filling in the reference's order.  With wd given, each linearised
    BC is applied to (w, wd) before the primal one, as handleBC.applyBC does
    (handleBC.py:184-240); for the flat plate the primal BCs were already
    applied, so the tangent call alone is what BROADCAST_npz.py:1079-1082 runs."""
    im, jm, gh = case.im, case.jm, case.gh
    gam = case.phys.gam
    it = case.interf
    if case.kind == "bl":
        lib.bc_supandsubinlet_2d(w, "Ilo", it["Ilo"], case.field, g["nx"], g["ny"], gam, im, jm,
                                 gh, wd=wd)
        lib.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], g["nx"], g["ny"], gam, gh, im, jm,
                               wd=wd)
        lib.bc_extrapolate_o2_2d(w, "Ihi", it["Ihi"], im, jm, gh, wd=wd)
        lib.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm, wd=wd)
    else:
        lib.bc_no_reflexion_2d(w, case.wbd, "Jhi", it["Jhi"], g["nx"], g["ny"], gam, gh, im, jm,
                               wd=wd)
        lib.bc_wall_viscous_adia_2d(w, "Jlo", gam, it["Jlo"], gh, im, jm, wd=wd)
        (prr1, prd2, tr), (prr2, prd1, _) = join_boxes(im, jm, gh)
        for _ in range(2):  # interf3 and interf4 both apply the two joins
            for arr in ([wd] if wd is not None else []) + [w]:
                lib.jn_match_2d(arr, prr1, gh, gh, gh, gh, im, jm, arr, prd2, gh, gh, gh, gh, im,
                                jm, tr)
                lib.jn_match_2d(arr, prr2, gh, gh, gh, gh, im, jm, arr, prd1, gh, gh, gh, gh, im,
                                jm, tr)


def _line_cells(loc, interf):
    """Cells (i, j) of the lines of a BC interface and its inward normal
    (i0, j0) (srcfv/borders/init_2d.F:1-44)."""
    b = np.asarray(interf).astype(int)
    (imin, jmin), (imax, jmax) = b[0], b[1]
    i0 = {"Ilo": 1, "Ihi": -1}.get(loc, 0)
    j0 = {"Jlo": 1, "Jhi": -1}.get(loc, 0)
    if i0:
        return [(imin, j) for j in range(jmin, jmax + 1)], i0, j0
    return [(i, jmin) for i in range(imin, imax + 1)], i0, j0


def adjoint_line_bc(case: Case, apply_d, w, wb, loc, interf):
    """Adjoint of one line BC (the reference's bc_*_b, srcfv/adjoint/bc_*_b.f90)
    formed as the transpose of its tangent: a BC line reads the cells at
    offsets 0..3 along its inward normal and writes its gh ghost layers, so
    for each offset d and variable q a unit tangent on that cell of every
    line (lines are disjoint) gives the column d ghosts / d w(d, q), and
    wb(d, q) += sum over the ghosts of column * wb(ghost).  Then wb(ghosts) = 0.
    apply_d(w, wd) applies the tangent BC in place.  Returns the new wb."""
    gh = case.gh
    cells, i0, j0 = _line_cells(loc, interf)
    I = lambda i: i + gh - 1
    out = np.array(wb, order="F", copy=True)
    ghosts = [[(I(i - de * i0), I(j - de * j0)) for de in range(1, gh + 1)] for i, j in cells]
    for d in range(4):
        for q in range(5):
            wc = np.array(w, order="F", copy=True)
            wd = np.zeros_like(wc, order="F")
            for i, j in cells:
                wd[I(i + d * i0), I(j + d * j0), q] = 1.0
            apply_d(wc, wd)
            for (i, j), gs in zip(cells, ghosts):
                acc = sum(float(np.dot(wd[a, b, :], wb[a, b, :])) for a, b in gs)
                out[I(i + d * i0), I(j + d * j0), q] += acc
    for gs in ghosts:
        for a, b in gs:
            out[a, b, :] = 0.0
    return out


def adjoint_join(case: Case, wb, prr, prd):
    """In-block join adjoint: the copy wb(prr) <- wb(prd) transposed."""
    gh = case.gh
    r = np.asarray(prr).astype(int)
    d = np.asarray(prd).astype(int)
    out = np.array(wb, order="F", copy=True)
    rs = (slice(r[0, 0] + gh - 1, r[1, 0] + gh), slice(r[0, 1] + gh - 1, r[1, 1] + gh))
    ds = (slice(d[0, 0] + gh - 1, d[1, 0] + gh), slice(d[0, 1] + gh - 1, d[1, 1] + gh))
    out[ds] += out[rs]
    out[rs] = 0.0
    return out


def adjoint_bcs(lib, case: Case, g, w, wb):
    """Adjoint of the BC sequence of apply_bcs, ops in reverse order, at the
    state w the sequence produced."""
    im, jm, gh = case.im, case.jm, case.gh
    gam = case.phys.gam
    it = case.interf
    ops = []
    if case.kind == "bl":
        ops.append(("Ilo", lambda a, ad: lib.bc_supandsubinlet_2d(a, "Ilo", it["Ilo"], case.field, g["nx"],
                                                                  g["ny"], gam, im, jm, gh, wd=ad)))
        ops.append(("Jhi", lambda a, ad: lib.bc_no_reflexion_2d(a, case.wbd, "Jhi", it["Jhi"], g["nx"], g["ny"],
                                                                gam, gh, im, jm, wd=ad)))
        ops.append(("Ihi", lambda a, ad: lib.bc_extrapolate_o2_2d(a, "Ihi", it["Ihi"], im, jm, gh, wd=ad)))
        ops.append(("Jlo", lambda a, ad: lib.bc_wall_viscous_adia_2d(a, "Jlo", gam, it["Jlo"], gh, im, jm, wd=ad)))
        for loc, fn in reversed(ops):
            wb = adjoint_line_bc(case, fn, w, wb, loc, it[loc])
        return wb
    ops.append(("Jhi", lambda a, ad: lib.bc_no_reflexion_2d(a, case.wbd, "Jhi", it["Jhi"], g["nx"], g["ny"],
                                                            gam, gh, im, jm, wd=ad)))
    ops.append(("Jlo", lambda a, ad: lib.bc_wall_viscous_adia_2d(a, "Jlo", gam, it["Jlo"], gh, im, jm, wd=ad)))
    (prr1, prd2, _), (prr2, prd1, _) = join_boxes(im, jm, gh)
    for _ in range(2):
        for a, b in ((prr2, prd1), (prr1, prd2)):
            wb = adjoint_join(case, wb, a, b)
    for loc, fn in reversed(ops):
        wb = adjoint_line_bc(case, fn, w, wb, loc, it[loc])
    return wb


def residual(lib, case: Case, g, w):
    res = np.zeros_like(w, order="F")
    lib.flux_num_dnc5_2d(res, w, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"],
                         g["vol"], g["volf"], case.gh, *case.phys.scheme_args(), case.im,
                         case.jm)
    return res


def jvp_flux(lib, case: Case, g, w, wd):
    """dR/dw . wd of the flux routine alone: wd's ghost frame as given."""
    res = np.zeros_like(w, order="F")
    resd = np.zeros_like(w, order="F")
    lib.flux_num_dnc5_2d_d(res, resd, np.asfortranarray(w), np.asfortranarray(wd), g["x0"], g["y0"],
                           g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], case.gh,
                           *case.phys.scheme_args(), case.im, case.jm)
    return resd


def vjp(lib, case: Case, g, w, resb):
    """wb = (dR/dw)^T resb of the flux routine alone (cylinder.py:1141);
    resb is not modified (a copy is consumed)."""
    rb = np.array(resb, order="F", copy=True)
    wb = np.zeros_like(rb, order="F")
    res = np.zeros_like(rb, order="F")
    lib.flux_num_dnc5_2d_b(res, rb, np.asfortranarray(w), wb, g["x0"], g["y0"], g["nx"], g["ny"],
                           g["xc"], g["yc"], g["vol"], g["volf"], case.gh, *case.phys.scheme_args(),
                           case.im, case.jm)
    return wb


def jvp(lib, case: Case, g, w, wd):
    """dR/dw . wd for interior wd: linearised BCs fill wd's ghosts, then the
    tangent residual (BROADCAST_npz.py:1079-1085)."""
    w = np.array(w, order="F", copy=True)
    wd = np.array(wd, order="F", copy=True)
    apply_bcs(lib, case, g, w, wd=wd)
    res = np.zeros_like(w, order="F")
    resd = np.zeros_like(w, order="F")
    lib.flux_num_dnc5_2d_d(res, resd, w, wd, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"],
                           g["yc"], g["vol"], g["volf"], case.gh, *case.phys.scheme_args(),
                           case.im, case.jm)
    return res, resd


def two_zone_split(case: Case):
    """First i index (0-based) of the upper zone when the reference colors an
    i-periodic mesh in two zones, None otherwise: an O-mesh whose im is not a
    multiple of 2gh+1 is split at (im/2)/(2gh+1)*(2gh+1) (cylinder.py:896-897)."""
    step = 2 * case.gh + 1
    if case.kind == "cyl" and case.im % step != 0:
        return (case.im // 2) // step * step
    return None


def colored_jacobian(lib, case: Case, g, w, coefdiag=None, prune=2e-16, coo=False):
    """The reference's colored assembly (BROADCAST_npz.py:1068-1137;
    cylinder.py:895-985 for the O-mesh): one-hot seeds with period 2gh+1,
    tangent BCs + JVP per seed, scatter into COO, drop |J| <= prune, then
    COO -> CSR summing duplicates (misc/PETSc_func.py:71-95).  An O-mesh
    with im not a multiple of 2gh+1 uses the two-zone seeds and scatter
    (testvector_partial, computejacobianfromjv_relaxed_withjnandcheck) over a
    doubled COO.  Returns a scipy.sparse.csr_matrix of -dR/dw (+ coefdiag on
    the diagonal), or the raw COO (jac, ia, ja) when coo=True."""
    import scipy.sparse as sp

    im, jm, gh = case.im, case.jm, case.gh
    step = 2 * gh + 1
    split = two_zone_split(case)
    nzones = 1 if split is None else 2
    nb = im * jm * step * step * 25 * nzones
    jac = np.zeros(nb)
    ia = np.zeros(nb, dtype=np.int32)
    ja = np.zeros(nb, dtype=np.int32)
    withjn = case.kind == "cyl"
    wd = np.zeros_like(w, order="F")
    zones = [(0, im - 1)] if split is None else [(0, split - 1), (split, im - 1)]
    for n, (istart, iend) in enumerate(zones):
        for m in range(5):
            for l in range(step):
                for k in range(step):
                    if split is None:
                        lib.testvector(wd, m, l, k, gh, im, jm)
                    else:
                        lib.testvector_partial(wd, m, l, k, gh, im, jm, istart, iend, 0, jm - 1)
                    _, resd = jvp(lib, case, g, w, wd)
                    if split is None:
                        lib.computejacobianfromjv(jac, ia, ja, resd, m, l, k, gh, im, jm,
                                                  coefdiag=coefdiag, withjn=withjn)
                    else:
                        cd = coefdiag if coefdiag is not None else np.zeros((im, jm), order="F")
                        lib.computejacobianfromjv_relaxed_withjnandcheck(jac, ia, ja, resd, m, l, k, gh, im,
                                                                         jm, cd, 2e-16, n)
    if coo:
        return jac, ia, ja
    keep = np.abs(jac) > prune
    n = 5 * im * jm
    return sp.csr_matrix((jac[keep], (ia[keep], ja[keep])), shape=(n, n))


def brute_jacobian(lib, case: Case, g, w):
    """Column-by-column -dR/dw with unit seeds on every interior unknown
    (5 im jm JVPs); independent of any coloring, used to pin the colored
    assembly on tiny meshes.  Rows/cols in the C-order ravel of (im, jm, 5)."""
    im, jm, gh = case.im, case.jm, case.gh
    n = 5 * im * jm
    J = np.zeros((n, n))
    for i in range(im):
        for j in range(jm):
            for m in range(5):
                wd = np.zeros_like(w, order="F")
                wd[gh + i, gh + j, m] = 1.0
                _, resd = jvp(lib, case, g, w, wd)
                col = (i * jm + j) * 5 + m
                J[:, col] = -np.ravel(resd[gh:gh + im, gh:gh + jm, :])
    return J


def dz_operators(lib, case: Case, g, w, wd):
    """coeffs_5p_dz and coeffs_5p_dz2 (srcfv/dz) of wd after the linearised
    BCs (BROADCAST_npz.py:1236-1243)."""
    w = np.array(w, order="F", copy=True)
    wd = np.array(wd, order="F", copy=True)
    apply_bcs(lib, case, g, w, wd=wd)
    dz = np.zeros_like(w, order="F")
    dz2 = np.zeros_like(w, order="F")
    a = case.phys.scheme_args()[:9]
    for fn, out in ((lib.coeffs_5p_dz, dz), (lib.coeffs_5p_dz2, dz2)):
        fn(out, w, wd, g["x0"], g["y0"], g["nx"], g["ny"], g["xc"], g["yc"], g["vol"], g["volf"], case.gh, *a,
           case.im, case.jm)
    return dz, dz2


def dz_colored(lib, case: Case, g, w):
    """The Dz/Dzz seed loop of BROADCAST_npz.py:1230-1245: raw COOs
    (jdz, jdz2, ia, ja) before remove_zero_jac."""
    im, jm, gh = case.im, case.jm, case.gh
    nb = im * jm * 49 * 25
    jdz, jdz2 = np.zeros(nb), np.zeros(nb)
    ia, ja = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
    ia2, ja2 = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
    wd = np.zeros_like(w, order="F")
    for m in range(5):
        for l in range(7):
            for k in range(7):
                lib.testvector(wd, m, l, k, gh, im, jm)
                dz, dz2 = dz_operators(lib, case, g, w, wd)
                lib.computejacobianfromdz(jdz, ia, ja, dz, m, l, k, gh, im, jm)
                lib.computejacobianfromdz(jdz2, ia2, ja2, dz2, m, l, k, gh, im, jm)
    assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2)
    return jdz, jdz2, ia, ja
