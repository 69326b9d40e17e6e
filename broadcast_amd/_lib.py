"""ctypes binding of libbroadcast_amd.so (the C ABI of include/broadcast_amd.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing (build it with
``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbroadcast_amd.so")

_dp = C.c_void_p


class Phys(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("cp", "cv", "prandtl", "gam", "rgaz", "cs", "muref", "tref", "s_suth", "k2", "k4")]


class Metrics(C.Structure):
    _fields_ = [("nx", _dp), ("ny", _dp), ("vol", _dp), ("volf", _dp)]


class BcOp(C.Structure):
    _fields_ = [("kind", C.c_int), ("loc", C.c_int), ("box", C.c_int * 4), ("box2", C.c_int * 4),
                ("tr", C.c_int * 2), ("data", _dp), ("lm", C.c_int), ("gam", C.c_double)]


BC_INLET, BC_NOREF, BC_EXTRAP2, BC_WALL_ADIA, BC_JOIN = range(5)


class BatchItem(C.Structure):
    _fields_ = [("im", C.c_int), ("jm", C.c_int), ("g", Metrics), ("phys", Phys), ("ops", C.POINTER(BcOp)),
                ("nops", C.c_int), ("w", _dp), ("wd", _dp), ("residu", _dp), ("residud", _dp)]

_SIGS = {
    "bca_ctx_create": [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, _dp],
    "bca_ctx_destroy": [_dp],
    "bca_ctx_set_stream": [_dp, _dp],
    "bca_ctx_scratch_bytes": [_dp],
    "bca_profile_enable": [_dp, C.c_int],
    "bca_profile_read": [_dp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)],
    "bca_geom": [_dp] * 9,
    "bca_volf": [_dp] * 3,
    "bca_residual_dnc5": [_dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_residual_dnc5_noalloc": [_dp, _dp, Metrics, C.POINTER(Phys), C.c_int, C.c_int, C.c_int, _dp],
    "bca_residual_noalloc_bytes": [C.c_int, C.c_int, C.c_int],
    "bca_jvp_dnc5": [_dp, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_residual_dnc5_nowall": [_dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_jvp_dnc5_nowall": [_dp, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_residual_dnc3": [_dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_residual_dnc3_nowall": [_dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_jvp_dnc3": [_dp, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_jvp_dnc3_nowall": [_dp, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_residual_dnc": [_dp, C.c_int, C.c_int, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_jvp_dnc": [_dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_vjp_dnc5": [_dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_bc_inlet": [_dp, _dp, _dp, C.c_int, C.POINTER(C.c_int), _dp, C.c_int, _dp, _dp, C.c_double],
    "bca_bc_noref": [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, C.c_double],
    "bca_bc_extrap2": [_dp, _dp, _dp, C.c_int, C.POINTER(C.c_int), C.c_int],
    "bca_bc_wall_adia": [_dp, _dp, _dp, C.c_int, C.POINTER(C.c_int), C.c_double],
    "bca_bc_join": [_dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                    C.c_int, C.c_int],
    "bca_norml2inf": [_dp, _dp, C.POINTER(C.c_double)],
    "bca_norml2": [_dp, _dp, C.POINTER(C.c_double)],
    "bca_seed": [_dp, _dp, C.c_int, C.c_int, C.c_int],
    "bca_scatter_jv": [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_long],
    "bca_seed_partial": [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "bca_scatter_jv_2zones": [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_int,
                              C.c_long],
    "bca_set_bndbl": [_dp, _dp, _dp, _dp],
    "bca_jacobian_colored": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, _dp,
                             C.c_int, _dp, _dp, _dp, C.c_long, C.c_int],
    "bca_csr_colored_count": [_dp, _dp, C.c_long, C.c_int, C.c_int, C.c_double, _dp, C.POINTER(C.c_long)],
    "bca_csr_colored_fill": [_dp, _dp, C.c_long, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp],
    "bca_dz_5p": [_dp, _dp, _dp, _dp, _dp, Metrics, C.POINTER(Phys)],
    "bca_scatter_dz": [_dp, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_long],
    "bca_dz_colored": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, _dp, _dp, _dp, _dp,
                       C.c_long, C.c_int],
    "bca_jbeta_count": [_dp, _dp, _dp, _dp, C.c_long, C.c_double, _dp, C.POINTER(C.c_long)],
    "bca_jbeta_fill": [_dp, _dp, _dp, _dp, C.c_long, C.c_double, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_long],
    "bca_jacobian_colored_2zones": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, _dp,
                                    C.c_int, _dp, _dp, _dp, C.c_long, C.c_int],
    "bca_batch_create": [C.POINTER(C.c_void_p), _dp],
    "bca_batch_set_stream": [_dp, _dp],
    "bca_batch_destroy": [_dp],
    "bca_batch_residual_jvp": [_dp, C.POINTER(BatchItem), C.c_int],
    "bca_bc_adjoint": [_dp, _dp, _dp, C.POINTER(BcOp), _dp, _dp],
    "bca_jacobian_colored_csr": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, _dp, C.c_int,
                                 C.c_int, C.c_double, _dp, _dp, _dp, C.c_long, C.POINTER(C.c_long)],
    "bca_jacobian_colored_csr_refresh": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, _dp,
                                         C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_long,
                                         C.POINTER(C.c_long), C.POINTER(C.c_int)],
    "bca_autotune": [_dp, _dp, Metrics, C.POINTER(Phys), C.POINTER(BcOp), C.c_int, C.c_int, C.c_int, _dp, C.c_long,
                     C.c_int, C.c_char_p, C.c_int],
    "bca_gemm_bf16": [_dp, C.c_int, C.c_int, C.c_int, _dp, C.c_long, C.c_long, _dp, C.c_long, C.c_long, _dp,
                      C.c_long, C.c_long, C.c_int, C.c_int],
    "bca_lu_np": [_dp, _dp, C.c_int, C.c_long, C.c_double, _dp, _dp],
    "bca_inv_np": [_dp, _dp, C.c_int, C.c_long, _dp, C.c_double, _dp],
    "bca_gemm_f64_sub": [_dp, C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long, _dp, C.c_long],
    "bca_csr_spmv": [_dp, C.c_long, _dp, _dp, _dp, _dp, _dp],
}

_LONG_RESULT = ("bca_ctx_scratch_bytes", "bca_residual_noalloc_bytes")

# Every symbol include/broadcast_amd.h declares (checked by the CPU test suite).
EXPORTS = sorted(list(_SIGS) + ["bca_last_error"])

_lib = None


class BcaError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: the HIP library must be built (__graft_entry__.build()); "
                "broadcast_amd has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, argtypes in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = C.c_long if name in _LONG_RESULT else C.c_int
        L.bca_last_error.restype = C.c_char_p
        L.bca_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().bca_last_error().decode(errors="replace")
        raise BcaError(f"{what} failed ({rc}): {msg}")
