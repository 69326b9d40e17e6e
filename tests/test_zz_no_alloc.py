"""The residual entry point that allocates nothing (bca_residual_dnc5_noalloc,
scratch in one static device array) against bca_residual_dnc5 (context
scratch from hipMalloc), through the C ABI: bit for bit, on two calls in a
row, the second on a smaller mesh so that it runs over the first call's
leftovers in the shared array; and a mesh beyond the array refused before
any launch."""
import ctypes as C

import numpy as np
import pytest

from broadcast_amd.cases import bl_case, cylinder_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


def _both(t, case):
    from broadcast_amd import _lib
    from broadcast_amd.flow import Flow, to_device

    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    ref = f.residual(w)
    got = t.full_like(ref, 7.0)  # ghosts must stay as given, interior overwritten
    s = t.cuda.current_stream().cuda_stream
    rc = _lib.lib().bca_residual_dnc5_noalloc(C.c_void_p(got.data_ptr()), C.c_void_p(w.data_ptr()), f.metrics,
                                              C.byref(f.phys), case.im, case.jm, case.gh, C.c_void_p(s))
    _lib.check(rc, "bca_residual_dnc5_noalloc")
    t.cuda.synchronize()
    ni, nj = case.im + 6, case.jm + 6
    a = got.view(5, nj, ni)
    b = ref.view(5, nj, ni)
    assert t.equal(a[:, 3:-3, 3:-3], b[:, 3:-3, 3:-3])
    ghost = t.ones_like(a, dtype=t.bool)
    ghost[:, 3:-3, 3:-3] = False
    assert bool((a[ghost] == 7.0).all())
    return f


def test_noalloc_residual_matches_context_path_twice(torch_mod):
    t = torch_mod
    _both(t, cylinder_case(96, 64, seed=5))
    _both(t, bl_case(40, 24, seed=9))   # smaller, over the first call's leftovers


def test_noalloc_residual_refuses_oversized_mesh(torch_mod):
    from broadcast_amd import _lib

    L = _lib.lib()
    need = L.bca_residual_noalloc_bytes(2048, 1024, 3)
    assert need > 64 * 1024 * 1024
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    m = _lib.Metrics(dummy, dummy, dummy, dummy)
    ph = _lib.Phys()
    rc = L.bca_residual_dnc5_noalloc(dummy, dummy, m, C.byref(ph), 2048, 1024, 3, None)
    assert rc == -4
    assert "scratch" in L.bca_last_error().decode()
