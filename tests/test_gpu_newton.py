"""Newton driver on the device (broadcast_amd.newton): the reference's
fixed-point iteration from the seeded Blasius-type state with 1e-3 noise.
The reference card's cfl = 1e10 (card_bl2d_fv_npz.py:54-56, pure Newton from
a noise-free similarity state) does not converge from the noisy state (the
noise on 1e-4-high wall cells puts it outside Newton's basin); with the
reference's own residual-scaled pseudo-time term at cfl = 1e3 it does.  The iterate sequence depends on the linear
solver's rounding (the reference's MUMPS LU has no pinned version), so the
driver is pinned by its end state: max_e ||R_e||_2 < 1e-10, and the
converged state's residual evaluated by the oracle agrees."""
import numpy as np
import pytest

import oracle
from oracle import driver as D
from broadcast_amd.cases import bl_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    assert torch.cuda.is_available(), "no HIP device"
    return torch


@pytest.mark.parametrize("mk", [lambda: bl_case(48, 32, seed=61), lambda: bl_case(256, 128, seed=0)])
def test_newton_converges_and_oracle_agrees(torch_mod, mk):
    from broadcast_amd.flow import Flow, to_device, to_host
    from broadcast_amd.newton import newton

    case = mk()
    print(f"case bl {case.im}x{case.jm}", flush=True)
    f = Flow(case)
    w = to_device(case.w)
    log = newton(f, w, ite=40, cfl=1e3, tol=1e-10, verbose=True)
    print(f"  norms {['%.2e' % v for v in log.norms]}  seconds {['%.2f' % v for v in log.seconds]}", flush=True)
    assert log.converged, log.norms
    assert log.rollbacks == 0
    O = oracle.Oracle()
    g = D.geometry(O, case)
    wh = np.array(to_host(w, f.shape_w), order="F", copy=True)
    D.apply_bcs(O, case, g, wh)
    res = D.residual(O, case, g, wh)
    norm, _ = O.compute_norml2inf(res, case.im, case.jm, case.gh)
    assert np.max(norm) < 1e-10 * 1.05, norm
