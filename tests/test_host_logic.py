"""Host-side logic that runs without a GPU: parameter sampling and scaling of
the learned guess, the uneven-mesh case generator, the exported C symbols of
the new entry points."""
import numpy as np
import pytest


def test_sobol_params_range_and_determinism():
    from broadcast_amd.learned import MACH_RANGE, RE_RANGE, sobol_params

    p = sobol_params(112, seed=4)
    assert p.shape == (112, 2)
    assert p[:, 0].min() >= RE_RANGE[0] and p[:, 0].max() <= RE_RANGE[1]
    assert p[:, 1].min() >= MACH_RANGE[0] and p[:, 1].max() <= MACH_RANGE[1]
    np.testing.assert_array_equal(p, sobol_params(112, seed=4))
    # skip continues the same sequence
    np.testing.assert_array_equal(sobol_params(16, seed=4, skip=96), p[96:])
    # a low-discrepancy set covers the square: every quarter holds points
    q = (p - [RE_RANGE[0], MACH_RANGE[0]]) / [RE_RANGE[1] - RE_RANGE[0], MACH_RANGE[1] - MACH_RANGE[0]]
    for a in (0, 1):
        for b in (0, 1):
            assert np.sum((q[:, 0] >= a / 2) & (q[:, 0] < (a + 1) / 2) & (q[:, 1] >= b / 2) & (q[:, 1] < (b + 1) / 2)) >= 20


def test_scale_params_maps_the_box_to_unit_square():
    from broadcast_amd.learned import MACH_RANGE, RE_RANGE, scale_params

    corners = np.array([[RE_RANGE[0], MACH_RANGE[0]], [RE_RANGE[1], MACH_RANGE[1]], [110.0, 0.25]])
    s = scale_params(corners)
    np.testing.assert_allclose(s[0], [-1, -1])
    np.testing.assert_allclose(s[1], [1, 1])
    np.testing.assert_allclose(s[2], [0, 0], atol=1e-15)


def test_uneven_cases_log_uniform_in_the_stated_ranges():
    from broadcast_amd.cases import uneven_bl_cases

    cases = uneven_bl_cases(8, seed=2000)
    assert len(cases) == 8
    for c in cases:
        assert 128 <= c.im <= 1024 and 64 <= c.jm <= 256
        assert c.w.shape == (c.im + 6, c.jm + 6, 5)
    # seeded: the same sizes every time
    assert [(c.im, c.jm) for c in cases] == [(c.im, c.jm) for c in uneven_bl_cases(8, seed=2000)]


def test_new_entry_points_are_exported():
    import os

    from broadcast_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    for name in ("bca_batch_create", "bca_batch_residual_jvp", "bca_bc_adjoint", "bca_jacobian_colored_csr",
                 "bca_gemm_bf16", "bca_autotune"):
        assert hasattr(L, name), name
