"""Synthetic inputs for the FP64 residual / JVP path.

Two mesh families are produced, both as node coordinates plus a cell state in
the reference's storage convention (Fortran order, ghost width ``gh`` on every
side, node arrays of shape ``(im+2gh+1, jm+2gh+1)`` and cell arrays of shape
``(im+2gh, jm+2gh[, 5])``):

* ``bl_case``  -- 2-D laminar flat-plate boundary layer on one structured
  block (the setup of BROADCAST_npz.py:474-799): x uniform, y geometrically
  stretched, inflow at Ilo, extrapolation at Ihi, adiabatic no-slip wall at
  Jlo, non-reflecting far field at Jhi, with the interface boxes of
  BROADCAST_npz.py:706-731.
* ``cylinder_case`` -- O-mesh around a circular cylinder (cylinder.py:431-530):
  periodic join in i, wall at Jlo, non-reflecting far field at Jhi.

Non-dimensionalisation: rho_inf = U_inf = T_inf = 1, so that
rgaz = 1/(gam M^2), cv = rgaz/(gam-1), cp = gam cv and Sutherland's mu_ref is
chosen so that mu(T_inf) = 1/Re.

Geometry (normals, volumes) is NOT computed here: that is the job of
``f_geom.computegeom_2d`` (the product's HIP kernel) or of the oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dc_field
import math

import numpy as np

GH = 3  # ghost width for the 5th-order DNC scheme: gh = (order+1)//2


@dataclass
class Physics:
    mach: float
    reynolds: float
    gam: float = 1.4
    prandtl: float = 0.72
    t_suth_kelvin: float = 110.4
    t_inf_kelvin: float = 288.15
    k2: float = 1.01
    k4: float = 1.0

    @property
    def rgaz(self) -> float:
        return 1.0 / (self.gam * self.mach * self.mach)

    @property
    def cv(self) -> float:
        return self.rgaz / (self.gam - 1.0)

    @property
    def cp(self) -> float:
        return self.gam * self.cv

    @property
    def cs(self) -> float:
        """Sutherland constant S / T_inf (non-dimensional)."""
        return self.t_suth_kelvin / self.t_inf_kelvin

    @property
    def tref(self) -> float:
        return 1.0

    @property
    def muref(self) -> float:
        # mu(T) = betas T^1.5/(T+S), betas = muref (Tref+S)/Tref^1.5, so mu(Tref) = muref
        return 1.0 / self.reynolds

    @property
    def p_inf(self) -> float:
        return self.rgaz  # rho R T with rho = T = 1

    @property
    def rhoe_inf(self) -> float:
        return self.p_inf / (self.gam - 1.0) + 0.5

    def state_inf(self) -> np.ndarray:
        return np.array([1.0, 1.0, 0.0, 0.0, self.rhoe_inf])

    def scheme_args(self):
        """(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4) in the
        positional order of flux_num_dnc5_2d (srcfv/rhs/flux_num_dnc5.F90:7-8);
        the reference driver passes cs twice, as cs and as s_suth
        (BROADCAST_npz.py:1031)."""
        return (self.cp, self.cv, self.prandtl, self.gam, self.rgaz, self.cs,
                self.muref, self.tref, self.cs, self.k2, self.k4)


@dataclass
class Case:
    name: str
    kind: str  # "bl" or "cyl"
    im: int
    jm: int
    gh: int
    phys: Physics
    x0: np.ndarray  # (im+2gh+1, jm+2gh+1) F-order, interior nodes filled
    y0: np.ndarray
    w: np.ndarray  # (im+2gh, jm+2gh, 5) F-order
    field: np.ndarray | None = None  # inflow ghost profile (jm, gh, 5)
    wbd: np.ndarray | None = None  # far-field state (lm, 5)
    interf: dict = dc_field(default_factory=dict)  # name -> (2,2) float64 F-order

    @property
    def ni(self) -> int:
        return self.im + 2 * self.gh

    @property
    def nj(self) -> int:
        return self.jm + 2 * self.gh


def _interf(imin, jmin, imax, jmax) -> np.ndarray:
    a = np.zeros((2, 2), order="F")
    a[0, 0], a[0, 1], a[1, 0], a[1, 1] = imin, jmin, imax, jmax
    return a


def geometric_nodes(n: int, d1: float, ratio: float) -> np.ndarray:
    """n+1 node coordinates starting at 0 with first spacing d1 and growth ratio."""
    d = d1 * ratio ** np.arange(n)
    return np.concatenate(([0.0], np.cumsum(d)))


def _blasius_table(eta_max: float = 15.0, n: int = 15001):
    """Blasius f''' + f f''/2 = 0, f(0)=f'(0)=0, f'(inf)=1: RK4 shooting on
    s = f''(0), secant iteration on f'(eta_max; s) - 1 (smooth in s)."""
    h = eta_max / (n - 1)

    def integrate(s, keep=False):
        a, b, c = 0.0, 0.0, s
        out = np.empty((n, 3)) if keep else None
        if keep:
            out[0] = (a, b, c)
        for k in range(1, n):
            k1a, k1b, k1c = b, c, -0.5 * a * c
            a2, b2, c2 = a + 0.5 * h * k1a, b + 0.5 * h * k1b, c + 0.5 * h * k1c
            k2a, k2b, k2c = b2, c2, -0.5 * a2 * c2
            a3, b3, c3 = a + 0.5 * h * k2a, b + 0.5 * h * k2b, c + 0.5 * h * k2c
            k3a, k3b, k3c = b3, c3, -0.5 * a3 * c3
            a4, b4, c4 = a + h * k3a, b + h * k3b, c + h * k3c
            k4a, k4b, k4c = b4, c4, -0.5 * a4 * c4
            a += h / 6.0 * (k1a + 2 * k2a + 2 * k3a + k4a)
            b += h / 6.0 * (k1b + 2 * k2b + 2 * k3b + k4b)
            c += h / 6.0 * (k1c + 2 * k2c + 2 * k3c + k4c)
            if keep:
                out[k] = (a, b, c)
        return out if keep else b

    s0, s1 = 0.3, 0.35
    f0, f1 = integrate(s0) - 1.0, integrate(s1) - 1.0
    for _ in range(50):
        if f1 == f0:
            break
        s0, f0, s1 = s1, f1, s1 - f1 * (s1 - s0) / (f1 - f0)
        f1 = integrate(s1) - 1.0
        if abs(s1 - s0) <= 1e-15 * abs(s1):
            break
    sol = integrate(s1, keep=True)
    eta = np.linspace(0.0, eta_max, n)
    return eta, sol[:, 0], sol[:, 1]


_BLASIUS = None


def blasius_state(xc, yc, phys: Physics):
    """Compressible Blasius-type state at cell centres (xc > 0 measured from
    the leading edge): u/U = f'(eta), eta = y sqrt(Re/x), Crocco-Busemann
    adiabatic-wall temperature with recovery factor sqrt(Pr), rho = 1/T,
    v = (eta f' - f)/(2 sqrt(Re x))."""
    global _BLASIUS
    if _BLASIUS is None:
        _BLASIUS = _blasius_table()
    eta_t, f_t, fp_t = _BLASIUS
    x = np.maximum(xc, 1e-12)
    y = np.maximum(yc, 0.0)
    eta = y * np.sqrt(phys.reynolds / x)
    fp = np.interp(eta, eta_t, fp_t, right=1.0)
    f = np.interp(eta, eta_t, f_t, right=f_t[-1] + (eta.max() - eta_t[-1]))
    f = np.where(eta > eta_t[-1], f_t[-1] + (eta - eta_t[-1]), f)
    u = fp
    v = (eta * fp - f) / (2.0 * np.sqrt(phys.reynolds * x))
    r = math.sqrt(phys.prandtl)
    t = 1.0 + r * 0.5 * (phys.gam - 1.0) * phys.mach ** 2 * (1.0 - u * u)
    rho = 1.0 / t
    p = phys.p_inf
    rhoe = p / (phys.gam - 1.0) + 0.5 * rho * (u * u + v * v)
    return rho, rho * u, rho * v, np.zeros_like(rho), rhoe


def bl_case(im: int, jm: int, *, mach=0.5, reynolds=1e5, x_le=0.1, length=1.0,
            dy1=1e-4, ratio=1.03, noise=1e-3, seed=0, k2=1.01, k4=1.0,
            name=None, gh=GH) -> Case:
    """Flat-plate boundary layer (BROADCAST_npz.py:474-799 set-up, own mesh and
    initial state).  Interior nodes: x uniform on [x_le, x_le+length], y
    geometric from 0 with first spacing dy1 and ratio.  State: Blasius-type +
    N(0, (noise*q_inf)^2) on the interior (q_inf = (1,1,1,1,rhoE_inf)).
    gh = 2 gives the ghost width of the 3rd-order scheme."""
    phys = Physics(mach=mach, reynolds=reynolds, k2=k2, k4=k4)
    ni, nj = im + 2 * gh, jm + 2 * gh
    xn = np.linspace(x_le, x_le + length, im + 1)
    yn = geometric_nodes(jm, dy1, ratio)
    x0 = np.zeros((ni + 1, nj + 1), order="F")
    y0 = np.zeros((ni + 1, nj + 1), order="F")
    x0[gh:gh + im + 1, :] = xn[:, None]
    y0[:, gh:gh + jm + 1] = yn[None, :]
    # cell centres (interior + ghost extrapolation only needed for the profile)
    xcn = 0.5 * (xn[1:] + xn[:-1])
    ycn = 0.5 * (yn[1:] + yn[:-1])
    dx = xn[1] - xn[0]
    xc_full = np.concatenate([xcn[0] - dx * np.arange(gh, 0, -1), xcn,
                              xcn[-1] + dx * np.arange(1, gh + 1)])
    yc_full = np.concatenate([-ycn[:gh][::-1], ycn,
                              ycn[-1] + (ycn[-1] - ycn[-2]) * np.arange(1, gh + 1)])
    # ghost columns upstream of a coarse inflow could reach x <= 0, where the
    # similarity profile is singular: evaluate them no closer than x_le/2.
    XC, YC = np.meshgrid(np.maximum(xc_full, 0.5 * x_le), yc_full, indexing="ij")
    comps = blasius_state(XC, YC, phys)
    w = np.zeros((ni, nj, 5), order="F")
    for e in range(5):
        w[:, :, e] = comps[e]
    # inflow ghost profile and far-field state from the clean profile, as the
    # reference's set_bndbl_2d does right after the similarity initialisation
    # (BROADCAST_npz.py:761-773, set_bnd.f90:2-24):
    # field(j,d,:) = w(1-d,j,:), wbd(i,:) = w(i-gh,jm,:)
    field = np.zeros((jm, gh, 5), order="F")
    for d in range(1, gh + 1):
        field[:, d - 1, :] = w[gh - d, gh:gh + jm, :]
    wbd = np.zeros((im + gh, 5), order="F")
    wbd[:, :] = w[0:im + gh, gh + jm - 1, :]
    rng = np.random.default_rng(seed)
    qinf = np.array([1.0, 1.0, 1.0, 1.0, phys.rhoe_inf])
    w[gh:gh + im, gh:gh + jm, :] += noise * qinf * rng.standard_normal((im, jm, 5))
    interf = {
        "Ilo": _interf(1, 1, 1, jm),
        "Ihi": _interf(im, 1, im, jm + gh),
        "Jlo": _interf(1 - gh, 1, im + gh, 1),
        "Jhi": _interf(1 - gh, jm, im, jm),
    }
    return Case(name or f"bl{im}x{jm}", "bl", im, jm, gh, phys, x0, y0, w,
                field, wbd, interf)


def uneven_bl_cases(n: int = 8, seed: int = 2000, im_range=(128, 1024), jm_range=(64, 256)):
    """Boundary-layer meshes of log-uniform random size: instance k draws
    im in im_range and jm in jm_range log-uniformly from seed + k and uses
    seed + k for its state noise too."""
    out = []
    for k in range(n):
        rng = np.random.default_rng(seed + k)
        im = int(round(np.exp(rng.uniform(np.log(im_range[0]), np.log(im_range[1])))))
        jm = int(round(np.exp(rng.uniform(np.log(jm_range[0]), np.log(jm_range[1])))))
        out.append(bl_case(im, jm, seed=seed + k))
    return out


def cylinder_case(im: int, jm: int, *, mach=0.1, reynolds=50.0, r0=0.5,
                  r1=100.0, dr1=5e-3, noise=1e-4, seed=0, k2=0.0, k4=1.0,
                  name=None, gh=GH) -> Case:
    """O-mesh around a cylinder of diameter 1 (cylinder.py:431-530 set-up, own
    mesh): i runs clockwise in theta (so that (i, j) is right-handed like the
    flat-plate block), j runs outward from r0 to r1 with first spacing dr1 and
    a geometric ratio q solving dr1 (q^jm - 1)/(q - 1) = r1 - r0.  State:
    uniform free stream along +x plus N(0, noise^2) on the interior.
    gh = 2 gives the ghost width of the 3rd-order scheme."""
    phys = Physics(mach=mach, reynolds=reynolds, k2=k2, k4=k4)
    ni, nj = im + 2 * gh, jm + 2 * gh
    span = r1 - r0
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        q = 0.5 * (lo + hi)
        s = dr1 * (q ** jm - 1.0) / (q - 1.0)
        if s > span:
            hi = q
        else:
            lo = q
    q = 0.5 * (lo + hi)
    rn = r0 + geometric_nodes(jm, dr1, q)
    rn[-1] = r1
    th = -2.0 * math.pi * np.arange(im + 1) / im
    x0 = np.zeros((ni + 1, nj + 1), order="F")
    y0 = np.zeros((ni + 1, nj + 1), order="F")
    x0[gh:gh + im + 1, gh:gh + jm + 1] = rn[None, :] * np.cos(th)[:, None]
    y0[gh:gh + im + 1, gh:gh + jm + 1] = rn[None, :] * np.sin(th)[:, None]
    w = np.zeros((ni, nj, 5), order="F")
    w[:, :, :] = phys.state_inf()[None, None, :]
    rng = np.random.default_rng(seed)
    w[gh:gh + im, gh:gh + jm, :] += noise * rng.standard_normal((im, jm, 5))
    wbd = np.zeros((max(im, jm) + gh, 5), order="F")
    wbd[:, :] = phys.state_inf()[None, :]
    interf = {
        "Jlo": _interf(1, 1, im, 1),
        "Jhi": _interf(1, jm, im, jm),
        "Ilo": _interf(1, 1 - gh, 1, jm + gh),
        "Ihi": _interf(im, 1 - gh, im, jm + gh),
    }
    return Case(name or f"cyl{im}x{jm}", "cyl", im, jm, gh, phys, x0, y0, w,
                None, wbd, interf)


def join_boxes(im: int, jm: int, gh: int):
    """Periodic-join boxes of cylinder.py:478-486: (prr1 <- prd2) fills the Ihi
    ghost columns from cells 1..gh, (prr2 <- prd1) fills the Ilo ghost columns
    from cells im-gh+1..im, over j = 1-gh..jm+gh.  Returned as int32 (2,2)
    F-order boxes [[imin, jmin], [imax, jmax]] plus the transform (1, 2)."""
    jmin, jmax = 1 - gh, jm + gh

    def box(a, b, c, d):
        return np.asfortranarray(np.array([[a, b], [c, d]], dtype=np.int32))

    prr1 = box(im + 1, jmin, im + gh, jmax)
    prd1 = box(im - gh + 1, jmin, im, jmax)
    prr2 = box(1 - gh, jmin, 0, jmax)
    prd2 = box(1, jmin, gh, jmax)
    tr = np.array([1, 2], dtype=np.int32)
    return (prr1, prd2, tr), (prr2, prd1, tr)
