"""Uneven-mesh batch (one launch per stage for the BCs + residual + JVP of
several meshes) against the per-mesh path and the oracle.

The per-mesh path (Flow.apply_bcs + bca_jvp_dnc5) is itself pinned against
the reference's golden vectors; the batch runs the same device functions in
another kernel, where the compiler may contract a different set of products
into FMAs, so the bounds are those of the per-mesh parity tests, normwise per
variable: max|a - b| <= tol max|b| with tol = 1e-14 for ghost states, 1e-12
for residuals and JVPs (the residual of a near-uniform flow is a difference of
O(1) face fluxes, so rounding shows at ~1e-13 of its size).  One small batch
is also checked against the oracle's BCs + tangent residual directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


def _rel(a, b):
    a = a.reshape(-1, 5)
    b = b.reshape(-1, 5)
    return np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)


def _inputs(torch_mod, flows, seed):
    from broadcast_amd.flow import to_device

    rng = np.random.default_rng(seed)
    ws, wds = [], []
    for f in flows:
        ws.append(to_device(f.case.w))
        wds.append(to_device(rng.standard_normal(f.shape_w)))
    return ws, wds


def _check_against_per_mesh(torch_mod, flows, tangent=True, seed=7):
    from broadcast_amd.batch import UnevenBatch

    ws, wds = _inputs(torch_mod, flows, seed)
    ref_w = [w.clone() for w in ws]
    ref_wd = [wd.clone() for wd in wds]
    b = UnevenBatch(flows)
    res, resd = b.run(ws, wds if tangent else None)
    for k, f in enumerate(flows):
        if tangent:
            r_ref, rd_ref = f.jvp(ref_w[k], ref_wd[k])
        else:
            f.apply_bcs(ref_w[k])
            r_ref = f.residual(ref_w[k])
        torch_mod.cuda.synchronize()
        v = lambda t: t.view(5, f.nj, f.ni).permute(2, 1, 0).cpu().numpy()
        inner = (slice(3, -3), slice(3, -3))
        assert np.all(_rel(v(ws[k]), v(ref_w[k])) <= 1e-14), f"mesh {k} ghost states"
        assert np.all(_rel(v(res[k])[inner], v(r_ref)[inner]) <= 1e-12), f"mesh {k} residual"
        if tangent:
            assert np.all(_rel(v(wds[k]), v(ref_wd[k])) <= 1e-14), f"mesh {k} ghost tangents"
            assert np.all(_rel(v(resd[k])[inner], v(rd_ref)[inner]) <= 1e-12), f"mesh {k} jvp"


def test_small_uneven_batch_with_omesh(torch_mod):
    from broadcast_amd.cases import bl_case, cylinder_case
    from broadcast_amd.flow import Flow

    cases = [bl_case(24, 16, seed=1), cylinder_case(28, 12, seed=2), bl_case(13, 9, seed=3),
             bl_case(40, 21, seed=4), cylinder_case(35, 10, seed=5)]
    flows = [Flow(c) for c in cases]
    _check_against_per_mesh(torch_mod, flows, tangent=True)
    _check_against_per_mesh(torch_mod, flows, tangent=False)


def test_batch_matches_oracle(torch_mod):
    import oracle
    from oracle import driver as D
    from broadcast_amd.batch import UnevenBatch
    from broadcast_amd.cases import bl_case, cylinder_case
    from broadcast_amd.flow import Flow, to_device, to_host
    from golden_util import close

    O = oracle.Oracle()
    cases = [bl_case(17, 11, seed=11), cylinder_case(21, 9, seed=12)]
    flows = [Flow(c) for c in cases]
    rng = np.random.default_rng(13)
    wd_h = [rng.standard_normal(f.shape_w) for f in flows]
    ws = [to_device(c.w) for c in cases]
    wds = [to_device(a) for a in wd_h]
    res, resd = UnevenBatch(flows).run(ws, wds)
    for k, (c, f) in enumerate(zip(cases, flows)):
        g = D.geometry(O, c)
        w_ref = np.array(c.w, order="F", copy=True)
        wd_ref = np.array(wd_h[k], order="F", copy=True)
        D.apply_bcs(O, c, g, w_ref, wd=wd_ref)
        r_ref = D.residual(O, c, g, w_ref)
        rd_ref = D.jvp_flux(O, c, g, w_ref, wd_ref)
        inner = (slice(3, -3), slice(3, -3))
        # bounds and scaling of the per-mesh parity tests (momentum components
        # share one scale: the BCs rotate momentum into normal/tangential parts)
        for got, ref, tol, what in ((ws[k], w_ref, 1e-14, "ghost states"), (wds[k], wd_ref, 1e-13, "ghost tangents"),
                                    (res[k], r_ref, 1e-12, "residual"), (resd[k], rd_ref, 1e-12, "jvp")):
            a = to_host(got, f.shape_w)
            if what in ("residual", "jvp"):
                a, ref = a[inner], ref[inner]
            ok, err = close(a, ref, tol)
            assert ok, f"mesh {k} {what}: {err}"


def test_full_size_uneven_batch(torch_mod):
    """The 8 log-uniform boundary-layer meshes (im 128..1024, jm 64..256)."""
    from broadcast_amd.cases import uneven_bl_cases
    from broadcast_amd.flow import Flow

    flows = [Flow(c) for c in uneven_bl_cases()]
    _check_against_per_mesh(torch_mod, flows, tangent=True, seed=2100)


def test_batch_errors(torch_mod):
    from broadcast_amd import _lib
    from broadcast_amd.batch import UnevenBatch
    from broadcast_amd.cases import bl_case
    from broadcast_amd.flow import Flow

    f = Flow(bl_case(13, 9))
    with pytest.raises(ValueError):
        UnevenBatch([])
    with pytest.raises(ValueError):
        UnevenBatch([f] * 17)
    b = UnevenBatch([f, f])
    w = [f.new_state(), f.new_state()]
    with pytest.raises(ValueError):
        b.run(w[:1])
    with pytest.raises(ValueError):
        b.run([w[0], torch_mod.zeros(10, dtype=torch_mod.float64, device="cuda")])
