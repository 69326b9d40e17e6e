"""The reference's NPZ "tree": the restart/exchange file between its Newton
driver and its stability tools (BROADCAST_npz.py:247-262, 311-333, 407-467).

Same file, same keys, same dtypes and shapes, so a tree written here from
device results is read unchanged by the reference's readNPZtree and by the
scripts that build PETSc matrices from its Jacobian lists (biglobal.py:40-61,
resolvent_all3D_1block.py), and a tree the reference wrote restarts a Flow:

* writeNPZ    -- geometry, initial state and residual, interface boxes, the BC
                 and scheme routine names ("f_bnd.bc_...", "f_sch.flux_...")
                 and the physical constants (:247-262);
* fillNPZ     -- adds the end-of-run state and residual and the Jacobian lists
                 IA, JA, Aij, with Aij the relaxed Jacobian divided by the row
                 cell volume (:311-320, :1204-1215);
* fillNPZ_3D  -- adds the pruned Dz / Dzz lists IAdz, JAdz, Aijdz, IAdz2,
                 JAdz2, Aijdz2 (:323-333, :1217-1253);
* readNPZtree -- the tuple the reference's reader returns, end-of-run state
                 first when present (:407-467).

Arrays may be numpy arrays or HIP torch tensors (copied to the host in the
reference's Fortran order).  `jacobian_lists` / `dz_lists` produce the lists
from a Flow's device COO in the reference's order: its remove_zero_jac
(:129-135) keeps the surviving COO slots in slot order, and so does the
order-preserving compaction here.
"""
from __future__ import annotations

import numpy as np

# the reference's routine names of the flat-plate case, in the order of its
# list lf = [inflow, outflow, far field, wall, scheme, libbnd, libsch]
# (card_bl2d_fv_npz.py:90-98)
BL_ROUTINES = ["bc_supandsubinlet_2d", "bc_extrapolate_o2_2d", "bc_no_reflexion_2d",
               "bc_wall_viscous_adia_2d", "flux_num_dnc5_2d", "f_bnd", "f_sch"]


def _host(a, shape=None):
    """numpy view of a numpy array or a device tensor (flat device states
    reshaped to `shape` in Fortran order)."""
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
        if shape is not None:
            a = a.reshape(shape, order="F")
    return np.asarray(a)


def _path(filename):
    return filename + ".npz"


def _str(a):
    """The reference reads a stored string as np.array2string(a)[1:-1]."""
    return np.array2string(np.asarray(a))[1:-1]


def writeNPZ(filename, im, jm, gh, w, x0, y0, vol, volf, nx, ny, xc, yc, field, wbd, res, sch, k2, k4,
             interf1, interf2, interf3, interf4, lf, cp, cv, gam, cs, tref, muref, rgaz, mach, prandtl, cfl,
             freqres, freqsort, pinf=None):
    """BROADCAST_npz.py:247-262, same positional arguments and keys."""
    ni, nj = im + 2 * gh, jm + 2 * gh
    w0 = np.array(_host(w, (ni, nj, 5)), copy=True)
    res0 = np.array(_host(res, (ni, nj, 5)), copy=True)
    routinein, routineout, routinenr, routinew, routinesch, libbnd, libsch = lf
    g = {k: _host(v, s) for k, v, s in (("x0", x0, (ni + 1, nj + 1)), ("y0", y0, (ni + 1, nj + 1)),
                                       ("xc", xc, (ni, nj)), ("yc", yc, (ni, nj)), ("vol", vol, (ni, nj)),
                                       ("volf", volf, (ni, nj, 2)), ("nx", nx, (ni + 1, nj + 1, 2)),
                                       ("ny", ny, (ni + 1, nj + 1, 2)))}
    np.savez(_path(filename), **g, im=im, jm=jm, gh=gh, FlowSolutionInit=w0, ResidualInit=res0,
             interf1=interf1, interf2=interf2, interf3=interf3, interf4=interf4,
             BC1=libbnd + "." + routinein, BC2=libbnd + "." + routineout, BC3=libbnd + "." + routinew,
             BC4=libbnd + "." + routinenr, Wbd=_host(wbd), Field=_host(field), SchemeType=sch,
             Scheme=libsch + "." + routinesch, k2=k2, k4=k4, Mach=mach, Prandtl=prandtl, Rgaz=rgaz, Cp=cp, Cv=cv,
             Gamma=gam, TemperatureSutherland=cs, TemperatureRefSutherland=tref, ViscosityRefSutherland=muref,
             Pref=pinf, CFL=cfl, Freqres=freqres, Freqsort=freqsort)


def _update(filename, **keys):
    dic = dict(np.load(_path(filename), allow_pickle=True))
    dic.update(keys)
    np.savez(_path(filename), **dic)


def fillNPZ(filename, w, res, IA, JA, Jacvol, gh):
    """BROADCAST_npz.py:311-320: end-of-run state and residual, Jacobian lists."""
    dic = dict(np.load(_path(filename), allow_pickle=True))
    ni, nj = int(dic["im"]) + 2 * gh, int(dic["jm"]) + 2 * gh
    _update(filename, ResidualEndOfRun=_host(res, (ni, nj, 5)), FlowSolutionEndOfRun=_host(w, (ni, nj, 5)),
            IA=_host(IA), JA=_host(JA), Aij=_host(Jacvol))


def fillNPZ_3D(filename, IAdz, JAdz, Jacdz, IAdz2, JAdz2, Jacdz2):
    """BROADCAST_npz.py:323-333: the Dz and Dzz lists."""
    _update(filename, IAdz=_host(IAdz), JAdz=_host(JAdz), Aijdz=_host(Jacdz), IAdz2=_host(IAdz2),
            JAdz2=_host(JAdz2), Aijdz2=_host(Jacdz2))


def readNPZtree(filename):
    """BROADCAST_npz.py:407-467: the same 36-tuple (the end-of-run state and
    residual when the tree holds them, else the initial ones)."""
    dic = np.load(_path(filename), allow_pickle=True)
    im, jm, gh = dic["im"], dic["jm"], dic["gh"]
    if "FlowSolutionEndOfRun" in dic:
        w, res = dic["FlowSolutionEndOfRun"], dic["ResidualEndOfRun"]
    else:
        w, res = dic["FlowSolutionInit"], dic["ResidualInit"]
    libsch, routinesch = _str(dic["Scheme"]).split(".")
    libbnd, routinein = _str(dic["BC1"]).split(".")
    libbnd, routineout = _str(dic["BC2"]).split(".")
    libbnd, routinew = _str(dic["BC3"]).split(".")
    libbnd, routinenr = _str(dic["BC4"]).split(".")
    lf = [routinein, routineout, routinenr, routinew, routinesch, libbnd, libsch]
    return (im, jm, gh, w, dic["x0"], dic["y0"], dic["vol"], dic["volf"], dic["nx"], dic["ny"], dic["xc"],
            dic["yc"], dic["Field"], dic["Wbd"], res, _str(dic["SchemeType"]), dic["k2"], dic["k4"],
            dic["interf1"], dic["interf2"], dic["interf3"], dic["interf4"], lf, dic["Cp"], dic["Cv"],
            dic["Gamma"], dic["TemperatureSutherland"], dic["TemperatureRefSutherland"],
            dic["ViscosityRefSutherland"], dic["Rgaz"], dic["Mach"], dic["Prandtl"], dic["Pref"], dic["CFL"],
            dic["Freqres"], dic["Freqsort"])


# ------------------------------------------------------------------ device lists
def remove_zero_jac(IA, JA, Jac, mini=2e-16):
    """BROADCAST_npz.py:129-135 on device (or host) arrays: the entries with
    |Jac| > mini, in their COO order."""
    keep = abs(Jac) > mini
    return IA[keep], JA[keep], Jac[keep]


def jacobian_lists(flow, jac, ia, ja, mini=2e-16):
    """(IA, JA, Jacvol) of the reference's post-Newton step from a Flow's
    colored COO (BROADCAST_npz.py:1137, 1204-1209): pruned in slot order,
    each value divided by the volume of its row's cell."""
    IA, JA, J = remove_zero_jac(ia, ja, jac, mini)
    t = flow.t
    gh = flow.gh
    vol = flow.vol.view(flow.nj, flow.ni)                       # Fortran (ni, nj) storage
    row = IA.long()
    i = row // (flow.jm * 5) + gh
    j = (row % (flow.jm * 5)) // 5 + gh
    return IA, JA, J / vol[j, i].to(t.float64)


def dz_lists(jdz, jdz2, ia, ja, mini=2e-16):
    """(IAdz, JAdz, Aijdz, IAdz2, JAdz2, Aijdz2): the Dz and Dzz COOs of
    Flow.dz_colored, each pruned in slot order (BROADCAST_npz.py:1248-1249)."""
    a = remove_zero_jac(ia, ja, jdz, mini)
    b = remove_zero_jac(ia, ja, jdz2, mini)
    return a + b


def flow_lf(case):
    """The reference's routine-name list lf of a case (the names its
    readNPZtree resolves by string, BROADCAST_npz.py:825-829)."""
    if case.kind != "bl":
        raise ValueError("the NPZ tree of the reference's npz driver describes the flat-plate case")
    return list(BL_ROUTINES)


def case_from_tree(filename):
    """A broadcast_amd Case from a tree (the reference's bl2d_fromNPZtree
    inputs, BROADCAST_npz.py:802-829): mesh nodes, the stored state, the BC
    data and interface boxes, and the scheme constants.  The BC routines
    must be the flat-plate sequence the Flow runs (inflow, far field,
    outflow, wall)."""
    from .cases import Case

    (im, jm, gh, w, x0, y0, vol, volf, nx, ny, xc, yc, field, wbd, res, sch, k2, k4, i1, i2, i3, i4, lf, cp, cv,
     gam, cs, tref, muref, rgaz, mach, prandtl, pinf, cfl, freqres, freqsort) = readNPZtree(filename)
    if lf[:4] != BL_ROUTINES[:4] or lf[4] != "flux_num_dnc5_2d":
        raise ValueError(f"tree routines {lf} are not the flat-plate DNC5 sequence")
    im, jm, gh = int(im), int(jm), int(gh)
    phys = _TreePhys(tuple(float(v) for v in (cp, cv, prandtl, gam, rgaz, cs, muref, tref, cs, k2, k4)),
                     float(mach), float(gam))
    return Case(name=f"npz:{filename}", kind="bl", im=im, jm=jm, gh=gh, phys=phys,
                x0=np.asfortranarray(x0), y0=np.asfortranarray(y0), w=np.asfortranarray(w),
                field=np.asfortranarray(field), wbd=np.asfortranarray(wbd),
                interf={"Ilo": np.asfortranarray(i1), "Ihi": np.asfortranarray(i2),
                        "Jlo": np.asfortranarray(i3), "Jhi": np.asfortranarray(i4)})


class _TreePhys:
    """Scheme constants as stored in a tree (the positional scheme arguments
    cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)."""

    def __init__(self, args, mach, gam):
        self._args = args
        self.mach = mach
        self.gam = gam

    def scheme_args(self):
        return self._args
