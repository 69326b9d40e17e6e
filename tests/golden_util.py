"""Loading of the golden fixtures into cases the checkers can run."""
import os

import numpy as np

from broadcast_amd.cases import Case, Physics

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))


def case_from(fx):
    ph = fx["phys"]
    phys = Physics(mach=1.0, reynolds=1.0)  # scheme arguments are overridden below
    kind = str(fx["kind"])
    case = Case(name=kind, kind=kind, im=int(fx["im"]), jm=int(fx["jm"]), gh=int(fx["gh"]),
                phys=phys, x0=np.asfortranarray(fx["x0_in"]), y0=np.asfortranarray(fx["y0_in"]),
                w=np.asfortranarray(fx["w_in"]),
                field=np.asfortranarray(fx["field"]) if fx["field"].size else None,
                wbd=np.asfortranarray(fx["wbd"]),
                interf={k: np.asfortranarray(fx[f"interf_{k}"]) for k in ("Ilo", "Ihi", "Jlo", "Jhi")})
    # keep the exact scheme arguments the fixture was produced with
    args = tuple(float(v) for v in ph)
    case.phys = _FixedPhys(args, phys.gam if len(args) < 4 else args[3])
    return case


class _FixedPhys:
    def __init__(self, args, gam):
        self._args = args
        self.gam = gam

    def scheme_args(self):
        return self._args


def close(a, b, rtol, scale=None):
    """Normwise-scaled comparison: |a-b| <= rtol * max|b| per variable (last
    axis).  For 5-variable states the three momentum components share one
    scale, max |rho u_k| over k: boundary conditions rotate momentum into
    normal and tangential parts, so a small component inherits the rounding of
    the large one."""
    a = np.asarray(a)
    b = np.asarray(b)
    if scale is None:
        scale = np.max(np.abs(b).reshape(-1, b.shape[-1]), axis=0) if b.ndim >= 2 else np.max(np.abs(b))
        if b.ndim >= 2 and b.shape[-1] == 5:
            scale = scale.copy()
            scale[1:4] = scale[1:4].max()
    err = np.max(np.abs(a - b).reshape(-1, b.shape[-1]), axis=0) if b.ndim >= 2 else np.max(np.abs(a - b))
    return np.all(err <= rtol * np.maximum(scale, 1e-300)), err / np.maximum(scale, 1e-300)


def two_zone_diag_multiplicity(im, gh=3):
    """How many times the two-zone relaxed assembly of an O-mesh adds coefdiag
    to the diagonal of row cell i (1-based), per i: once per zone whose
    window puts the diagonal's color on column i itself.  Each zone's pass
    (computejacobianfromjv_relaxed_withjnandcheck, misc/ComputeJacobian.f90:
    1095-1204) writes coefdiag - resd into its own slot of the doubled COO
    whenever that slot's column is the row's own cell, and the CSR sums the
    two.  The window column of color l for zone n follows the reference's
    vali rules (:1121-1197); tests pin this count against the reference
    build's scatter on small meshes."""
    step = 2 * gh + 1
    mult = np.zeros(im + 1, dtype=int)
    for i in range(1, im + 1):
        for n in (0, 1):
            hit = False
            for l in range(step):
                if l <= gh:
                    v = l if i <= l + 1 + gh else (i - (l + 1) + gh) // step * step + l
                else:
                    v = l if i <= l - gh else (i - (l + 1) + gh) // step * step + l
                if n == 0:
                    if v >= im - 2 * gh:
                        v = l
                elif v <= 2 * gh:
                    v = im - im % step + l
                    if v > im - 1:
                        v -= step
                hit = hit or v == i - 1
            mult[i] += int(hit)
    return mult[1:]
