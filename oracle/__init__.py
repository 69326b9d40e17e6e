"""ORACLE -- test infrastructure only.

ctypes bindings for the two CPU checkers of the FP64 residual / tangent path:

* ``Oracle``    -- oracle/liboracle.so, the C++ restatement (oracle/bcoracle.cpp).
* ``Reference`` -- oracle/_ref/libbroadcast_ref.so, the reference Fortran built
  from /root/reference by oracle/Makefile.ref (present only where it was built).

Both expose the reference's f2py-style positional call forms
(BROADCAST_npz.py:1018-1031, 1072-1127) on numpy float64 F-order arrays,
mutating their outputs in place, so tests read like the reference's driver.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package.  Nothing in broadcast_amd/ imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libbroadcast_ref.so")

LOC = {"Ilo": 0, "Ihi": 1, "Jlo": 2, "Jhi": 3}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray)
    if a.dtype == np.float64:
        assert a.flags.f_contiguous, "arrays must be Fortran-contiguous"
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        assert a.flags.f_contiguous
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def _box(interf):
    a = np.asarray(interf)
    return np.array([int(a[0, 0]), int(a[0, 1]), int(a[1, 0]), int(a[1, 1])],
                    dtype=np.int32)


def _i2(interf):
    """(2,2) interface box as Fortran int32 (f2py casts the float64 box)."""
    return np.asfortranarray(np.asarray(interf).astype(np.int32))


def _run_big_stack(fn, stack_mb=1024):
    """Run fn() on a thread with a large stack: the reference Fortran keeps its
    scratch arrays as automatic (stack) arrays (srcfv/rhs/flux_num_dnc5.F90:29-37)."""
    out = {}

    def target():
        try:
            out["r"] = fn()
        except BaseException as e:  # pragma: no cover
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(stack_mb * 1024 * 1024)
    try:
        t = threading.Thread(target=target)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


class Oracle:
    """The C++ restatement (oracle/bcoracle.cpp)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = C.CDLL(path)

    @staticmethod
    def _phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4):
        return np.array([cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4],
                        dtype=np.float64)

    # f_geom ------------------------------------------------------------------
    def computegeom_2d(self, x0, y0, nx, ny, xc, yc, vol, volf, im, jm, gh):
        self.lib.or_computegeom_2d(_p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc),
                                   _p(vol), _p(volf), C.c_int(im), C.c_int(jm), C.c_int(gh))

    # f_sch / f_lin -----------------------------------------------------------
    def flux_num_dnc5_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv,
                         prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        ph = self._phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
        self.lib.or_flux_num_dnc5_2d(_p(res), _p(w), _p(nx), _p(ny), _p(vol), _p(volf),
                                     C.c_int(gh), _p(ph), C.c_int(im), C.c_int(jm))

    def flux_num_dnc5_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh,
                           cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        ph = self._phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
        self.lib.or_flux_num_dnc5_2d_d(_p(res), _p(resd), _p(w), _p(wd), _p(nx), _p(ny),
                                       _p(vol), _p(volf), C.c_int(gh), _p(ph),
                                       C.c_int(im), C.c_int(jm))

    def flux_num_dnc5_2d_b(self, res, resb, w, wb, x0, y0, nx, ny, xc, yc, vol, volf, gh,
                           cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        ph = self._phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4)
        self.lib.or_flux_num_dnc5_2d_b(_p(resb), _p(wb), _p(w), _p(nx), _p(ny), _p(vol), _p(volf),
                                       C.c_int(gh), _p(ph), C.c_int(im), C.c_int(jm))

    # f_bnd (primal) and f_lin (tangent, wd not None) --------------------------
    def _scheme(self, name, arrays, gh, phys, im, jm):
        args = [_p(a) for a in arrays] + [*self._i(gh), *self._r(*phys), *self._i(im, jm)]
        _run_big_stack(lambda: getattr(self.lib, name + "_")(*args))

    def scheme(self, order, nowall, tangent):
        """The reference's flux_num_dnc{order}[_nowall]_2d[_d] as a callable
        with the f2py argument order (srcfv/rhs, srcfv/tangent)."""
        name = f"flux_num_dnc{order}{'_nowall' if nowall else ''}_2d{'_d' if tangent else ''}"
        narr = 12 if tangent else 10

        def call(*args):
            arrays, gh, rest = args[:narr], args[narr], args[narr + 1:]
            self._scheme(name, arrays, gh, rest[:11], *rest[11:])
        return call

    def flux_num_dnc3_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/rhs/flux_num_dnc3.F90:7"""
        self._scheme("flux_num_dnc3_2d", (res, w, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_nowall_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/rhs/flux_num_dnc3_nowall.F90:7"""
        self._scheme("flux_num_dnc3_nowall_2d", (res, w, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/tangent/flux_num_dnc3_d.f90"""
        self._scheme("flux_num_dnc3_2d_d", (res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_nowall_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/tangent/flux_num_dnc3_nowall_d.f90"""
        self._scheme("flux_num_dnc3_nowall_2d_d", (res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def bc_supandsubinlet_2d(self, w, loc, interf, field, nx, ny, gam, im, jm, gh=3, wd=None):
        lm = field.shape[0]
        self.lib.or_bc_inlet(_p(w), _p(wd), C.c_int(LOC[loc]), _p(_box(interf)), _p(field),
                             _p(nx), _p(ny), C.c_double(gam), C.c_int(im), C.c_int(jm),
                             C.c_int(lm), C.c_int(gh))

    def bc_no_reflexion_2d(self, w, wbd, loc, interf, nx, ny, gam, gh, im, jm, wd=None):
        lm = wbd.shape[0]
        self.lib.or_bc_noref(_p(w), _p(wd), _p(wbd), C.c_int(LOC[loc]), _p(_box(interf)),
                             _p(nx), _p(ny), C.c_double(gam), C.c_int(gh), C.c_int(im),
                             C.c_int(jm), C.c_int(lm))

    def bc_extrapolate_o2_2d(self, w, loc, interf, im, jm, gh, wd=None):
        em = w.shape[2] if w.ndim == 3 else 1
        self.lib.or_bc_extrap2(_p(w), _p(wd), C.c_int(LOC[loc]), _p(_box(interf)), C.c_int(im),
                               C.c_int(jm), C.c_int(gh), C.c_int(em))

    def bc_wall_viscous_adia_2d(self, w, loc, gam, interf, gh, im, jm, wd=None):
        self.lib.or_bc_wall(_p(w), _p(wd), C.c_int(LOC[loc]), C.c_double(gam), _p(_box(interf)),
                            C.c_int(gh), C.c_int(im), C.c_int(jm))

    def jn_match_2d(self, wr, prr, gh1r, gh2r, gh3r, gh4r, imr, jmr, wd, prd, gh1d, gh2d,
                    gh3d, gh4d, imd, jmd, tr):
        em = wr.shape[2] if wr.ndim == 3 else 1
        self.lib.or_jn_match(_p(wr), _p(_i2(prr)), C.c_int(imr), C.c_int(jmr), _p(wd),
                             _p(_i2(prd)), C.c_int(imd), C.c_int(jmd),
                             _p(np.asarray(tr, dtype=np.int32)), C.c_int(gh1r), C.c_int(em))

    # f_norm ------------------------------------------------------------------
    def compute_norml2inf(self, res, im, jm, gh):
        n, ni = np.zeros(5), np.zeros(5)
        self.lib.or_compute_norml2inf(_p(n), _p(ni), _p(res), C.c_int(im), C.c_int(jm), C.c_int(gh))
        return n, ni

    def compute_norml2(self, res, im, jm, gh):
        n, nm = np.zeros(5), np.zeros(5)
        self.lib.or_compute_norml2(_p(n), _p(nm), _p(res), C.c_int(im), C.c_int(jm), C.c_int(gh))
        return n, nm

    # f_misc ------------------------------------------------------------------
    def testvector(self, wd, m, l, k, gh, im, jm):
        self.lib.or_testvector(_p(wd), C.c_int(m), C.c_int(l), C.c_int(k), C.c_int(gh),
                               C.c_int(im), C.c_int(jm))

    def computejacobianfromjv(self, jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag=None,
                              withjn=False):
        cd = None if coefdiag is None else np.asfortranarray(coefdiag, dtype=np.float64)
        self.lib.or_computejacobianfromjv(_p(jac), _p(ia), _p(ja), _p(resd), C.c_int(m),
                                          C.c_int(l), C.c_int(k), C.c_int(gh), C.c_int(im),
                                          C.c_int(jm), _p(cd), C.c_int(int(withjn)))

    def coeffs_5p_dz(self, dz, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz,
                     cs, muref, tref, s_suth, im, jm):
        ph = self._phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, 0.0, 0.0)
        self.lib.or_coeffs_5p_dz(_p(dz), _p(w), _p(wd), _p(nx), _p(ny), _p(vol), _p(volf), C.c_int(gh),
                                 _p(ph), C.c_int(im), C.c_int(jm))

    def coeffs_5p_dz2(self, dz2, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz,
                      cs, muref, tref, s_suth, im, jm):
        ph = self._phys(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, 0.0, 0.0)
        self.lib.or_coeffs_5p_dz2(_p(dz2), _p(w), _p(wd), C.c_int(gh), _p(ph), C.c_int(im), C.c_int(jm))

    def computejacobianfromdz(self, jac, ia, ja, dz, m, l, k, gh, im, jm):
        self.lib.or_computejacobianfromdz(_p(jac), _p(ia), _p(ja), _p(dz), *[C.c_int(v) for v in
                                                                             (m, l, k, gh, im, jm)])

    def testvector_partial(self, wd, m, l, k, gh, im, jm, istart, iend, jstart, jend):
        self.lib.or_testvector_partial(_p(wd), *[C.c_int(v) for v in (m, l, k, gh, im, jm, istart, iend,
                                                                      jstart, jend)])

    def computejacobianfromjv_relaxed_withjnandcheck(self, jac, ia, ja, resd, m, l, k, gh, im, jm,
                                                     coefdiag, mini, n):
        cd = None if coefdiag is None else np.asfortranarray(coefdiag, dtype=np.float64)
        self.lib.or_computejacobianfromjv_check(_p(jac), _p(ia), _p(ja), _p(resd), C.c_int(m), C.c_int(l),
                                                C.c_int(k), C.c_int(gh), C.c_int(im), C.c_int(jm), _p(cd),
                                                C.c_double(mini), C.c_int(n))

    def set_bndbl_2d(self, w, field, wbd, im, jm, gh):
        self.lib.or_set_bndbl_2d(_p(w), _p(field), _p(wbd), C.c_int(im), C.c_int(jm), C.c_int(gh))


class Reference:
    """The reference Fortran, called through its Fortran ABI (symbol_, every
    argument by reference, CHARACTER lengths appended as trailing size_t)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -f oracle/Makefile.ref)")
        self.lib = C.CDLL(path)

    @staticmethod
    def _r(*vals):
        return [C.byref(C.c_double(v)) for v in vals]

    @staticmethod
    def _i(*vals):
        return [C.byref(C.c_int(int(v))) for v in vals]

    def computegeom_2d(self, x0, y0, nx, ny, xc, yc, vol, volf, im, jm, gh):
        self.lib.computegeom_2d_(_p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc), _p(vol),
                                 _p(volf), *self._i(im, jm, gh))

    def flux_num_dnc5_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv,
                         prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        args = [_p(res), _p(w), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc), _p(vol),
                _p(volf), *self._i(gh),
                *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4),
                *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.flux_num_dnc5_2d_(*args))

    def flux_num_dnc5_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh,
                           cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        args = [_p(res), _p(resd), _p(w), _p(wd), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc),
                _p(yc), _p(vol), _p(volf), *self._i(gh),
                *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4),
                *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.flux_num_dnc5_2d_d_(*args))

    def flux_num_dnc5_2d_b(self, res, resb, w, wb, x0, y0, nx, ny, xc, yc, vol, volf, gh,
                           cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        args = [_p(res), _p(resb), _p(w), _p(wb), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc),
                _p(yc), _p(vol), _p(volf), *self._i(gh),
                *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4),
                *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.flux_num_dnc5_2d_b_(*args))

    def flux_num_dnc5_nowall_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv,
                                prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        args = [_p(res), _p(w), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc), _p(vol),
                _p(volf), *self._i(gh),
                *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4),
                *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.flux_num_dnc5_nowall_2d_(*args))

    def flux_num_dnc5_nowall_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh,
                                  cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4, im, jm):
        args = [_p(res), _p(resd), _p(w), _p(wd), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc),
                _p(yc), _p(vol), _p(volf), *self._i(gh),
                *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4),
                *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.flux_num_dnc5_nowall_2d_d_(*args))

    def _scheme(self, name, arrays, gh, phys, im, jm):
        args = [_p(a) for a in arrays] + [*self._i(gh), *self._r(*phys), *self._i(im, jm)]
        _run_big_stack(lambda: getattr(self.lib, name + "_")(*args))

    def scheme(self, order, nowall, tangent):
        """The reference's flux_num_dnc{order}[_nowall]_2d[_d] as a callable
        with the f2py argument order (srcfv/rhs, srcfv/tangent)."""
        name = f"flux_num_dnc{order}{'_nowall' if nowall else ''}_2d{'_d' if tangent else ''}"
        narr = 12 if tangent else 10

        def call(*args):
            arrays, gh, rest = args[:narr], args[narr], args[narr + 1:]
            self._scheme(name, arrays, gh, rest[:11], *rest[11:])
        return call

    def flux_num_dnc3_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/rhs/flux_num_dnc3.F90:7"""
        self._scheme("flux_num_dnc3_2d", (res, w, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_nowall_2d(self, res, w, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/rhs/flux_num_dnc3_nowall.F90:7"""
        self._scheme("flux_num_dnc3_nowall_2d", (res, w, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/tangent/flux_num_dnc3_d.f90"""
        self._scheme("flux_num_dnc3_2d_d", (res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def flux_num_dnc3_nowall_2d_d(self, res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, *phys_im_jm):
        """srcfv/tangent/flux_num_dnc3_nowall_d.f90"""
        self._scheme("flux_num_dnc3_nowall_2d_d", (res, resd, w, wd, x0, y0, nx, ny, xc, yc, vol, volf), gh,
                     phys_im_jm[:11], *phys_im_jm[11:])

    def bc_supandsubinlet_2d(self, w, loc, interf, field, nx, ny, gam, im, jm, gh=3, wd=None):
        lm = field.shape[0]
        if wd is None:
            self.lib.bc_supandsubinlet_2d_(_p(w), C.c_char_p(loc.encode()), _p(_i2(interf)),
                                           _p(field), _p(nx), _p(ny), *self._r(gam),
                                           *self._i(im, jm, lm, gh), C.c_size_t(3))
        else:
            self.lib.bc_supandsubinlet_2d_d_(_p(w), _p(wd), C.c_char_p(loc.encode()),
                                             _p(_i2(interf)), _p(field), _p(nx), _p(ny),
                                             *self._r(gam), *self._i(im, jm, lm, gh),
                                             C.c_size_t(3))

    def bc_no_reflexion_2d(self, w, wbd, loc, interf, nx, ny, gam, gh, im, jm, wd=None):
        lm = wbd.shape[0]
        if wd is None:
            self.lib.bc_no_reflexion_2d_(_p(w), _p(wbd), C.c_char_p(loc.encode()),
                                         _p(_i2(interf)), _p(nx), _p(ny), *self._r(gam),
                                         *self._i(gh, im, jm, lm), C.c_size_t(3))
        else:
            self.lib.bc_no_reflexion_2d_d_(_p(w), _p(wd), _p(wbd), C.c_char_p(loc.encode()),
                                           _p(_i2(interf)), _p(nx), _p(ny), *self._r(gam),
                                           *self._i(gh, im, jm, lm), C.c_size_t(3))

    def bc_extrapolate_o2_2d(self, w, loc, interf, im, jm, gh, wd=None):
        em = w.shape[2] if w.ndim == 3 else 1
        if wd is None:
            self.lib.bc_extrapolate_o2_2d_(_p(w), C.c_char_p(loc.encode()), _p(_i2(interf)),
                                           *self._i(im, jm, gh, em), C.c_size_t(3))
        else:
            self.lib.bc_extrapolate_o2_2d_d_(_p(w), _p(wd), C.c_char_p(loc.encode()),
                                             _p(_i2(interf)), *self._i(im, jm, gh, em),
                                             C.c_size_t(3))

    def bc_wall_viscous_adia_2d(self, w, loc, gam, interf, gh, im, jm, wd=None):
        if wd is None:
            self.lib.bc_wall_viscous_adia_2d_(_p(w), C.c_char_p(loc.encode()), *self._r(gam),
                                              _p(_i2(interf)), *self._i(gh, im, jm),
                                              C.c_size_t(3))
        else:
            self.lib.bc_wall_viscous_adia_2d_d_(_p(w), _p(wd), C.c_char_p(loc.encode()),
                                                *self._r(gam), _p(_i2(interf)),
                                                *self._i(gh, im, jm), C.c_size_t(3))

    # TAPENADE adjoints of the BCs (srcfv/adjoint/bc_*_b.f90): wb in/out, w restored
    def bc_supandsubinlet_2d_b(self, w, wb, loc, interf, field, nx, ny, gam, im, jm, gh=3):
        self.lib.bc_supandsubinlet_2d_b_(_p(w), _p(wb), C.c_char_p(loc.encode()), _p(_i2(interf)), _p(field),
                                         _p(nx), _p(ny), *self._r(gam), *self._i(im, jm, field.shape[0], gh),
                                         C.c_size_t(3))

    def bc_no_reflexion_2d_b(self, w, wb, wbd, loc, interf, nx, ny, gam, gh, im, jm):
        self.lib.bc_no_reflexion_2d_b_(_p(w), _p(wb), _p(wbd), C.c_char_p(loc.encode()), _p(_i2(interf)),
                                       _p(nx), _p(ny), *self._r(gam), *self._i(gh, im, jm, wbd.shape[0]),
                                       C.c_size_t(3))

    def bc_extrapolate_o2_2d_b(self, w, wb, loc, interf, im, jm, gh):
        self.lib.bc_extrapolate_o2_2d_b_(_p(w), _p(wb), C.c_char_p(loc.encode()), _p(_i2(interf)),
                                         *self._i(im, jm, gh, w.shape[2]), C.c_size_t(3))

    def bc_wall_viscous_adia_2d_b(self, w, wb, loc, gam, interf, gh, im, jm):
        self.lib.bc_wall_viscous_adia_2d_b_(_p(w), _p(wb), C.c_char_p(loc.encode()), *self._r(gam),
                                            _p(_i2(interf)), *self._i(gh, im, jm), C.c_size_t(3))

    def jn_match_2d(self, wr, prr, gh1r, gh2r, gh3r, gh4r, imr, jmr, wd, prd, gh1d, gh2d,
                    gh3d, gh4d, imd, jmd, tr):
        em = wr.shape[2] if wr.ndim == 3 else 1
        self.lib.jn_match_2d_(_p(wr), _p(_i2(prr)), *self._i(gh1r, gh2r, gh3r, gh4r, imr, jmr),
                              _p(wd), _p(_i2(prd)), *self._i(gh1d, gh2d, gh3d, gh4d, imd, jmd),
                              _p(np.asarray(tr, dtype=np.int32)), *self._i(em))

    def compute_norml2inf(self, res, im, jm, gh):
        n, ni = np.zeros(5), np.zeros(5)
        self.lib.compute_norml2inf_(_p(n), _p(ni), _p(res), *self._i(im, jm, gh))
        return n, ni

    def compute_norml2(self, res, im, jm, gh):
        n, nm = np.zeros(5), np.zeros(5)
        self.lib.compute_norml2_(_p(n), _p(nm), _p(res), *self._i(im, jm, gh))
        return n, nm

    def testvector(self, wd, m, l, k, gh, im, jm):
        self.lib.testvector_(_p(wd), *self._i(m, l, k, gh, im, jm))

    def computejacobianfromjv(self, jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag=None,
                              withjn=False):
        nb = jac.shape[0]
        if coefdiag is None:
            fn = self.lib.computejacobianfromjv_withjn_ if withjn else self.lib.computejacobianfromjv_
            fn(_p(jac), _p(ia), _p(ja), _p(resd), *self._i(m, l, k, gh, im, jm, nb))
        else:
            cd = np.asfortranarray(coefdiag, dtype=np.float64)
            fn = (self.lib.computejacobianfromjv_relaxed_withjn_ if withjn
                  else self.lib.computejacobianfromjv_relaxed_)
            fn(_p(jac), _p(ia), _p(ja), _p(resd), *self._i(m, l, k, gh, im, jm, nb), _p(cd))

    def coeffs_5p_dz(self, dz, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz,
                     cs, muref, tref, s_suth, im, jm):
        args = [_p(dz), _p(w), _p(wd), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc), _p(vol), _p(volf),
                *self._i(gh), *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth), *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.coeffs_5p_dz_(*args))

    def coeffs_5p_dz2(self, dz2, w, wd, x0, y0, nx, ny, xc, yc, vol, volf, gh, cp, cv, prandtl, gam, rgaz,
                      cs, muref, tref, s_suth, im, jm):
        args = [_p(dz2), _p(w), _p(wd), _p(x0), _p(y0), _p(nx), _p(ny), _p(xc), _p(yc), _p(vol), _p(volf),
                *self._i(gh), *self._r(cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth), *self._i(im, jm)]
        _run_big_stack(lambda: self.lib.coeffs_5p_dz2_(*args))

    def computejacobianfromdz(self, jac, ia, ja, dz, m, l, k, gh, im, jm):
        self.lib.computejacobianfromdz_(_p(jac), _p(ia), _p(ja), _p(dz), *self._i(m, l, k, gh, im, jm,
                                                                                  jac.shape[0]))

    def testvector_partial(self, wd, m, l, k, gh, im, jm, istart, iend, jstart, jend):
        self.lib.testvector_partial_(_p(wd), *self._i(m, l, k, gh, im, jm, istart, iend, jstart, jend))

    def computejacobianfromjv_relaxed_withjnandcheck(self, jac, ia, ja, resd, m, l, k, gh, im, jm,
                                                     coefdiag, mini, n):
        cd = np.asfortranarray(coefdiag, dtype=np.float64)
        self.lib.computejacobianfromjv_relaxed_withjnandcheck_(
            _p(jac), _p(ia), _p(ja), _p(resd), *self._i(m, l, k, gh, im, jm, jac.shape[0]), _p(cd),
            *self._r(mini), *self._i(n))

    def set_bndbl_2d(self, w, field, wbd, im, jm, gh):
        self.lib.set_bndbl_2d_(_p(w), _p(field), _p(wbd), *self._i(im, jm, gh))
