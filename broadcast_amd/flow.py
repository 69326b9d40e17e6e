"""Device-resident driver of the FP64 residual / JVP / Jacobian path.

``Flow`` keeps a case's geometry, boundary data and boundary-condition
sequence in HBM and runs the reference driver's per-iteration work on the
MI355X without host round trips: the BC sequence and residual of
BROADCAST_npz.py:1018-1031 (cylinder.py:853-868 for the O-mesh), the tangent
BCs + JVP of :1079-1085, and the whole 245-seed colored Jacobian assembly of
:1068-1127 as one call into the C ABI.

All device arrays are flat float64 tensors holding the reference's Fortran
layout (ghost frame included); ``to_device``/``to_host`` convert from/to numpy
F-order arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import device as dev
from .cases import Case, join_boxes


def to_device(a):
    t = dev.require_gpu()
    a = np.asfortranarray(a)
    return t.from_numpy(a.reshape(-1, order="F").copy()).to("cuda")


def to_host(t, shape):
    return np.asarray(t.detach().cpu().numpy()).reshape(shape, order="F")


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Flow:
    def __init__(self, case: Case):
        self.case = case
        self.t = dev.require_gpu()
        im, jm, gh = case.im, case.jm, case.gh
        self.im, self.jm, self.gh = im, jm, gh
        self.ni, self.nj = case.ni, case.nj
        self.shape_w = (self.ni, self.nj, 5)
        self.ctx = dev.context(im, jm, gh)
        self.L = _lib.lib()
        args = case.phys.scheme_args()
        self.phys = dev.phys(*args)
        self.gam = float(args[3])
        self._geometry()
        self._plan()

    # ------------------------------------------------------------ geometry
    def _geometry(self):
        c, t = self.case, self.t
        ni, nj = self.ni, self.nj
        z = lambda n: t.zeros(n, dtype=t.float64, device="cuda")
        self.x0 = to_device(c.x0)
        self.y0 = to_device(c.y0)
        self.nx = z((ni + 1) * (nj + 1) * 2)
        self.ny = z((ni + 1) * (nj + 1) * 2)
        self.xc, self.yc, self.vol = z(ni * nj), z(ni * nj), z(ni * nj)
        self.volf = z(ni * nj * 2)
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_geom(h, *map(_ptr, (self.x0, self.y0, self.nx, self.ny, self.xc, self.yc,
                                                  self.vol, self.volf))), "bca_geom")
        if c.kind == "cyl":
            # periodic ghost metrics (cylinder.py:496-527)
            im, jm, gh = self.im, self.jm, self.gh
            (prr1, prd2, tr), (prr2, prd1, _) = join_boxes(im, jm, gh)
            trc = (C.c_int * 2)(1, 2)
            for arr in (self.xc, self.yc, self.vol):
                for a, b in ((prr1, prd2), (prr2, prd1)):
                    _lib.check(self.L.bca_bc_join(h, _ptr(arr), dev.int4(a), dev.int4(b), trc, 1, ni, nj),
                               "bca_bc_join")
            jmin, jmax = 1 - gh, jm + gh
            box = lambda a, b, cc, d: np.array([[a, b], [cc, d]])
            prr1g, prd1g = box(im + 2, jmin, im + 1 + gh, jmax + 1), box(im + 1 - gh, jmin, im, jmax + 1)
            prr2g, prd2g = box(1 - gh, jmin, 0, jmax + 1), box(2, jmin, 1 + gh, jmax + 1)
            for arr in (self.nx, self.ny):
                for a, b in ((prr1g, prd2g), (prr2g, prd1g)):
                    _lib.check(self.L.bca_bc_join(h, _ptr(arr), dev.int4(a), dev.int4(b), trc, 2, ni + 1,
                                                  nj + 1), "bca_bc_join")
            _lib.check(self.L.bca_volf(h, _ptr(self.vol), _ptr(self.volf)), "bca_volf")
        self.metrics = _lib.Metrics(_ptr(self.nx), _ptr(self.ny), _ptr(self.vol), _ptr(self.volf))

    # ------------------------------------------------------------ BC sequence
    def _op(self, kind, loc=0, box=None, box2=None, data=None, lm=0):
        op = _lib.BcOp()
        op.kind, op.loc, op.gam = kind, loc, self.gam
        if box is not None:
            op.box[:] = list(dev.int4(box))
        if box2 is not None:
            op.box2[:] = list(dev.int4(box2))
        op.tr[:] = [1, 2]
        op.data = None if data is None else data.data_ptr()
        op.lm = lm
        return op

    def _plan(self):
        c = self.case
        it = c.interf
        ops = []
        if c.field is not None:
            self.field = to_device(c.field)
        self.wbd = to_device(c.wbd)
        if c.kind == "bl":
            # BROADCAST_npz.py:1018-1021: inflow, far field, outflow, wall
            ops.append(self._op(_lib.BC_INLET, 0, it["Ilo"], data=self.field, lm=c.field.shape[0]))
            ops.append(self._op(_lib.BC_NOREF, 3, it["Jhi"], data=self.wbd, lm=c.wbd.shape[0]))
            ops.append(self._op(_lib.BC_EXTRAP2, 1, it["Ihi"]))
            ops.append(self._op(_lib.BC_WALL_ADIA, 2, it["Jlo"]))
        else:
            # handleBC.sortBC order for the O-mesh: far field, wall, then the
            # periodic join (both directions, applied by each join interface)
            ops.append(self._op(_lib.BC_NOREF, 3, it["Jhi"], data=self.wbd, lm=c.wbd.shape[0]))
            ops.append(self._op(_lib.BC_WALL_ADIA, 2, it["Jlo"]))
            (prr1, prd2, _), (prr2, prd1, _) = join_boxes(self.im, self.jm, self.gh)
            for _ in range(2):
                ops.append(self._op(_lib.BC_JOIN, 0, prr1, prd2))
                ops.append(self._op(_lib.BC_JOIN, 0, prr2, prd1))
        self.ops = (_lib.BcOp * len(ops))(*ops)
        self.nops = len(ops)

    def apply_bcs(self, w, wd=None):
        """Ghost filling in place (w, and wd with the linearised BCs)."""
        h = self.ctx.sync_stream()
        L = self.L
        for op in self.ops:
            wp, wdp = _ptr(w), _ptr(wd)
            box = (C.c_int * 4)(*op.box)
            if op.kind == _lib.BC_INLET:
                rc = L.bca_bc_inlet(h, wp, wdp, op.loc, box, op.data, op.lm, _ptr(self.nx), _ptr(self.ny), op.gam)
            elif op.kind == _lib.BC_NOREF:
                rc = L.bca_bc_noref(h, wp, wdp, op.data, op.lm, op.loc, box, _ptr(self.nx), _ptr(self.ny), op.gam)
            elif op.kind == _lib.BC_EXTRAP2:
                rc = L.bca_bc_extrap2(h, wp, wdp, op.loc, box, 5)
            elif op.kind == _lib.BC_WALL_ADIA:
                rc = L.bca_bc_wall_adia(h, wp, wdp, op.loc, box, op.gam)
            else:
                box2 = (C.c_int * 4)(*op.box2)
                trc = (C.c_int * 2)(*op.tr)
                rc = 0
                for arr in ([wd] if wd is not None else []) + [w]:
                    rc = rc or L.bca_bc_join(h, _ptr(arr), box, box2, trc, 5, self.ni, self.nj)
            _lib.check(rc, "boundary condition")

    # ------------------------------------------------------------ operators
    def new_state(self):
        return self.t.zeros(self.ni * self.nj * 5, dtype=self.t.float64, device="cuda")

    def residual(self, w, res=None):
        res = self.new_state() if res is None else res
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_residual_dnc5(h, _ptr(res), _ptr(w), self.metrics, C.byref(self.phys)),
                   "bca_residual_dnc5")
        return res

    def jvp(self, w, wd, res=None, resd=None):
        """Tangent BCs on (w, wd) then res, resd = R(w), dR/dw . wd."""
        res = self.new_state() if res is None else res
        resd = self.new_state() if resd is None else resd
        self.apply_bcs(w, wd)
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_jvp_dnc5(h, _ptr(res), _ptr(resd), _ptr(w), _ptr(wd), self.metrics,
                                       C.byref(self.phys)), "bca_jvp_dnc5")
        return res, resd

    def jvp_flux(self, w, wd, res=None, resd=None):
        """dR/dw . wd of the flux routine alone (wd's ghost frame as given,
        srcfv/tangent/flux_num_dnc5_d.f90:15)."""
        res = self.new_state() if res is None else res
        resd = self.new_state() if resd is None else resd
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_jvp_dnc5(h, _ptr(res), _ptr(resd), _ptr(w), _ptr(wd), self.metrics,
                                       C.byref(self.phys)), "bca_jvp_dnc5")
        return res, resd

    def vjp(self, w, resb, wb=None, consume=False):
        """wb = (dR/dw)^T resb over the whole array, ghost frame included
        (srcfv/adjoint/flux_num_dnc5_b.f90:15).  The C call consumes resb's
        interior like the reference; unless consume=True a copy is passed."""
        wb = self.new_state() if wb is None else wb
        rb = resb if consume else resb.clone()
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_vjp_dnc5(h, _ptr(rb), _ptr(wb), _ptr(w), self.metrics, C.byref(self.phys)),
                   "bca_vjp_dnc5")
        return wb

    def adjoint_bcs(self, w, wb):
        """Adjoint of the BC sequence, in place on wb: the ops in reverse
        order, each pushing its ghost adjoints to the cells it reads and
        zeroing them (the reference's bc_*_b / jn_match_2d_b, cylinder.py:
        1121-1154).  w is the state the sequence produced (each BC then reads
        the inputs it read in the forward sequence: no BC here reads a cell
        that a later one writes)."""
        h = self.ctx.sync_stream()
        for op in reversed(list(self.ops)):
            _lib.check(self.L.bca_bc_adjoint(h, _ptr(w), _ptr(wb), C.byref(op), _ptr(self.nx), _ptr(self.ny)),
                       "bca_bc_adjoint")
        return wb

    def vjp_bcs(self, w, resb, wb=None):
        """(d (R o BC) / dw)^T resb: the flux adjoint, then the BC sequence's
        adjoint; the transpose of jvp(w, wd) as a map of wd."""
        wb = self.vjp(w, resb, wb)
        return self.adjoint_bcs(w, wb)

    def norms(self, res):
        out = (C.c_double * 10)()
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_norml2inf(h, _ptr(res), out), "bca_norml2inf")
        o = np.array(out[:])
        return o[:5], o[5:]

    @property
    def two_zone_split(self):
        """First i index (0-based) of the upper seed zone when the reference
        colors this mesh in two zones (an O-mesh whose im is not a multiple of
        2gh+1, cylinder.py:896-897), else None."""
        step = 2 * self.gh + 1
        if self.case.kind == "cyl" and self.im % step != 0:
            return (self.im // 2) // step * step
        return None

    @property
    def nbentry(self):
        s = 2 * self.gh + 1
        return self.im * self.jm * s * s * 25 * (1 if self.two_zone_split is None else 2)

    def jacobian_buffers(self):
        t = self.t
        n = self.nbentry
        return (t.zeros(n, dtype=t.float64, device="cuda"), t.zeros(n, dtype=t.int32, device="cuda"),
                t.zeros(n, dtype=t.int32, device="cuda"))

    def jacobian_colored(self, w, coefdiag=None, bufs=None, write_index=True):
        """The reference's colored assembly into its COO layout (jac, ia, ja)
        on the device: -dR/dw (+ coefdiag on the diagonal); w must carry its
        ghost frame (BCs applied).  Meshes the reference colors in two zones
        get the two-zone seeds and scatter over the doubled COO."""
        jac, ia, ja = self.jacobian_buffers() if bufs is None else bufs
        h = self.ctx.sync_stream()
        split = self.two_zone_split
        if split is None:
            rc = self.L.bca_jacobian_colored(h, _ptr(w), self.metrics, C.byref(self.phys), self.ops, self.nops,
                                             _ptr(coefdiag), int(self.case.kind == "cyl"), _ptr(jac), _ptr(ia),
                                             _ptr(ja), self.nbentry, int(write_index))
        else:
            rc = self.L.bca_jacobian_colored_2zones(h, _ptr(w), self.metrics, C.byref(self.phys), self.ops,
                                                    self.nops, _ptr(coefdiag), split, _ptr(jac), _ptr(ia),
                                                    _ptr(ja), self.nbentry, int(write_index))
        _lib.check(rc, "bca_jacobian_colored")
        return jac, ia, ja

    def csr(self, jac, mini=2e-16):
        """Pruned CSR (indptr, indices, data) of a COO from jacobian_colored:
        remove_zero_jac + csr_matrix (BROADCAST_npz.py:129-135,
        misc/PETSc_func.py:85) on the device."""
        t = self.t
        n = 5 * self.im * self.jm
        nz = 1 if self.two_zone_split is None else 2
        withjn = int(self.case.kind == "cyl")
        indptr = t.empty(n + 1, dtype=t.int32, device="cuda")
        nnz = C.c_long()
        h = self.ctx.sync_stream()
        if jac.numel() < self.nbentry:
            raise ValueError(f"jac holds {jac.numel()} values, the colored COO of this mesh {self.nbentry}")
        _lib.check(self.L.bca_csr_colored_count(h, _ptr(jac), jac.numel(), nz, withjn, float(mini), _ptr(indptr),
                                                C.byref(nnz)), "bca_csr_colored_count")
        indices = t.empty(max(nnz.value, 1), dtype=t.int32, device="cuda")
        data = t.empty(max(nnz.value, 1), dtype=t.float64, device="cuda")
        _lib.check(self.L.bca_csr_colored_fill(h, _ptr(jac), jac.numel(), nz, withjn, float(mini), _ptr(indptr),
                                               _ptr(indices), _ptr(data)), "bca_csr_colored_fill")
        return indptr, indices[:nnz.value], data[:nnz.value]

    def autotune(self, w, jac=None, reps=3):
        """Choose the launch shapes of the residual faces, the face blocks and
        the colored assembly for this mesh by timing candidates on w (ghosts
        as given); later calls on this Flow use them.  Returns the report."""
        split = self.two_zone_split
        nz, zp = (1, int(self.case.kind == "cyl")) if split is None else (2, split)
        scratch = jac if jac is not None else self.t.empty(self.nbentry, dtype=self.t.float64, device="cuda")
        buf = C.create_string_buffer(512)
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_autotune(h, _ptr(w), self.metrics, C.byref(self.phys), self.ops, self.nops, nz, zp,
                                       _ptr(scratch), scratch.numel(), int(reps), buf, 512), "bca_autotune")
        return buf.value.decode()

    def jacobian_csr(self, w, coefdiag=None, mini=2e-16, refresh=False):
        """The colored Jacobian straight into the pruned CSR
        (bca_jacobian_colored_csr): the (indptr, indices, data) that
        csr(jacobian_colored(w, coefdiag)) gives, without forming the COO.
        The value arrays are views of buffers kept for the next call (grown
        when a larger nnz needs it).  refresh=True (a Newton loop) goes
        through bca_jacobian_colored_csr_refresh: where no row's kept set
        changed since the previous call only data is rewritten; the result is
        the same.  self.csr_rebuilt tells whether indptr/indices were
        rewritten."""
        t = self.t
        n = 5 * self.im * self.jm
        split = self.two_zone_split
        nz, zp = (1, int(self.case.kind == "cyl")) if split is None else (2, split)
        bufs = getattr(self, "_csr_bufs", None)
        if bufs is None:
            # ~145 entries per interior row (29-cell stencil x 5 variables)
            cap = 160 * n
            bufs = (t.empty(n + 1, dtype=t.int32, device="cuda"), t.empty(cap, dtype=t.int32, device="cuda"),
                    t.empty(cap, dtype=t.float64, device="cuda"))
        nnz = C.c_long()
        for _ in range(2):
            indptr, indices, data = bufs
            h = self.ctx.sync_stream()
            if refresh:
                rebuilt = C.c_int(1)
                rc = self.L.bca_jacobian_colored_csr_refresh(
                    h, _ptr(w), self.metrics, C.byref(self.phys), self.ops, self.nops, _ptr(coefdiag), nz, zp,
                    float(mini), _ptr(indptr), _ptr(indices), _ptr(data), indices.numel(), C.byref(nnz),
                    C.byref(rebuilt))
                self.csr_rebuilt = bool(rebuilt.value)
            else:
                rc = self.L.bca_jacobian_colored_csr(h, _ptr(w), self.metrics, C.byref(self.phys), self.ops,
                                                     self.nops, _ptr(coefdiag), nz, zp, float(mini), _ptr(indptr),
                                                     _ptr(indices), _ptr(data), indices.numel(), C.byref(nnz))
                self.csr_rebuilt = True
            if rc != -5:
                break
            bufs = (indptr, t.empty(nnz.value, dtype=t.int32, device="cuda"),
                    t.empty(nnz.value, dtype=t.float64, device="cuda"))
        _lib.check(rc, "bca_jacobian_colored_csr")
        self._csr_bufs = bufs
        return bufs[0], bufs[1][:nnz.value], bufs[2][:nnz.value]

    # ------------------------------------------------------------ 3D operators
    def dz_colored(self, w, bufs=None, write_index=True):
        """Dz and Dzz of the seed loop of BROADCAST_npz.py:1230-1245 in their
        COO layout (jdz, jdz2, ia, ja), computejacobianfromdz windows; w must
        carry its ghost frame."""
        t = self.t
        s = 2 * self.gh + 1
        n = self.im * self.jm * s * s * 25
        if bufs is None:
            bufs = (t.zeros(n, dtype=t.float64, device="cuda"), t.zeros(n, dtype=t.float64, device="cuda"),
                    t.zeros(n, dtype=t.int32, device="cuda"), t.zeros(n, dtype=t.int32, device="cuda"))
        jdz, jdz2, ia, ja = bufs
        h = self.ctx.sync_stream()
        _lib.check(self.L.bca_dz_colored(h, _ptr(w), self.metrics, C.byref(self.phys), self.ops, self.nops,
                                         _ptr(jdz), _ptr(jdz2), _ptr(ia), _ptr(ja), n, int(write_index)),
                   "bca_dz_colored")
        return jdz, jdz2, ia, ja

    def jbeta(self, jac, jdz, jdz2, betas, mini=2e-16):
        """J(beta) = Jac/vol(row) - beta^2 Dzz + i beta Dz (BROADCAST_npz.py:1206-1209,
        biglobal.py:82) for every beta: (indptr, indices, values[nbeta, nnz]
        complex128) from the COOs of jacobian_colored and dz_colored."""
        if self.case.kind == "cyl":
            raise ValueError("J(beta) is composed on the single-zone boundary-layer coloring")
        t = self.t
        n = 5 * self.im * self.jm
        beta = t.as_tensor(list(map(float, betas)), dtype=t.float64, device="cuda")
        indptr = t.empty(n + 1, dtype=t.int32, device="cuda")
        nnz = C.c_long()
        h = self.ctx.sync_stream()
        nb = jac.numel()
        _lib.check(self.L.bca_jbeta_count(h, _ptr(jac), _ptr(jdz), _ptr(jdz2), nb, float(mini), _ptr(indptr),
                                          C.byref(nnz)), "bca_jbeta_count")
        indices = t.empty(max(nnz.value, 1), dtype=t.int32, device="cuda")
        vals = t.empty((beta.numel(), max(nnz.value, 1), 2), dtype=t.float64, device="cuda")
        _lib.check(self.L.bca_jbeta_fill(h, _ptr(jac), _ptr(jdz), _ptr(jdz2), nb, float(mini), _ptr(self.vol),
                                         _ptr(beta), beta.numel(), _ptr(indptr), _ptr(indices), _ptr(vals),
                                         nnz.value), "bca_jbeta_fill")
        return indptr, indices[:nnz.value], t.view_as_complex(vals)[:, :nnz.value]
