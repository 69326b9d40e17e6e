"""Learned Newton initial guess for the cylinder baseflow.

A small network maps the flow parameters (Re_D, M) to the FP64 conservative
field on the O-mesh; Newton (broadcast_amd.newton, the reference's
fixed_point iteration, cylinder.py:827-1045) then starts from that field
instead of the uniform free stream, and falls back to the free-stream start
when the guess does not converge.  No reference counterpart: the reference
always starts from its initial state.

Network (BF16 weights and activations on the matrix cores, FP32 accumulation,
FP32 master weights and Adam):
    x = (Re, M) scaled to [-1, 1]
    h1 = silu(x W1^T + b1), h2 = silu(h1 W2^T + b2)      (width H)
    c = h2 W3^T + b3                                      (K coefficients)
    y = c P + mu                                          (normalised field)
P (K x n, n = 5 im jm) is a trainable linear decoder initialised with the
leading POD modes of the training fields, so most of the parameters sit in it
(K n ~ 2M at 384 x 256 with K = 4, 250k at 96 x 64 with K = 8).  Every product,
forward and backward (X W^T, dY^T X, dY W), is bca_gemm_bf16, the project's
MFMA kernel; the elementwise parts (bias, silu, Adam) are torch glue.

Training data: baseflows the GPU Newton solver computes at Sobol-sampled
(Re, M) in [40, 180] x [0.1, 0.4], from the free stream.  The network learns
the baseflow's departure from the free stream of the same (Re, M), which is
known exactly (the energy alone varies as 1/M^2 over the range, the
departure is small and smooth): guess = free stream + prediction.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import device as dev

RE_RANGE = (40.0, 180.0)
MACH_RANGE = (0.1, 0.4)


# ------------------------------------------------------------------ sampling
def sobol_params(n, seed=4, skip=0):
    """n (Re, M) pairs of a scrambled 2-D Sobol sequence over RE_RANGE x
    MACH_RANGE (points skip .. skip + n - 1 of the sequence)."""
    from scipy.stats import qmc

    s = qmc.Sobol(d=2, scramble=True, seed=seed)
    total = 1 << max(1, int(math.ceil(math.log2(skip + n))))
    u = s.random(total)[skip:skip + n]
    lo = np.array([RE_RANGE[0], MACH_RANGE[0]])
    hi = np.array([RE_RANGE[1], MACH_RANGE[1]])
    return lo + u * (hi - lo)


def scale_params(p):
    """(Re, M) -> [-1, 1]^2."""
    p = np.asarray(p, dtype=np.float64)
    lo = np.array([RE_RANGE[0], MACH_RANGE[0]])
    hi = np.array([RE_RANGE[1], MACH_RANGE[1]])
    return 2.0 * (p - lo) / (hi - lo) - 1.0


# ------------------------------------------------------------------ GEMM
def mm(a, b, out=None, accumulate=False, ksplit=1):
    """out (fp32, M x N) = a (bf16, M x K) @ b (bf16, K x N) on the matrix
    cores (bca_gemm_bf16); a and b may be any strided views."""
    t = dev.torch()
    if a.dtype != t.bfloat16 or b.dtype != t.bfloat16:
        raise TypeError("mm takes bfloat16 operands")
    if a.dim() != 2 or b.dim() != 2 or a.shape[1] != b.shape[0]:
        raise ValueError(f"shapes {tuple(a.shape)} @ {tuple(b.shape)}")
    M, K = a.shape
    N = b.shape[1]
    if out is None:
        out = (t.zeros if (accumulate or ksplit > 1) else t.empty)((M, N), dtype=t.float32, device=a.device)
    L = _lib.lib()
    _lib.check(L.bca_gemm_bf16(dev.current_stream_ptr(), M, N, K, C.c_void_p(a.data_ptr()), a.stride(0),
                               a.stride(1), C.c_void_p(b.data_ptr()), b.stride(0), b.stride(1),
                               C.c_void_p(out.data_ptr()), out.stride(0), out.stride(1), int(accumulate),
                               int(ksplit)), "bca_gemm_bf16")
    return out


def _ksplit(K):
    return max(1, min(1024, K // 2048))


# ------------------------------------------------------------------ network
@dataclass
class GuessNet:
    n: int                    # field length 5 im jm
    width: int = 64
    modes: int = 8
    params: dict = field(default_factory=dict)    # fp32 master weights
    scale: object = None      # per-entry field scale (n,), fp32
    mean: object = None       # mean normalised field (n,), fp32

    @staticmethod
    def create(fields, width=64, modes=8, seed=0):
        """Network for training fields (N x n, FP64 device tensor): field
        normalisation, decoder initialised with the leading POD modes."""
        t = dev.torch()
        N, n = fields.shape
        net = GuessNet(n=n, width=width, modes=modes)
        f = fields.view(N, 5, -1)
        s = f.abs().amax(dim=(0, 2)).clamp_min(1e-12)                 # per variable
        net.scale = s.repeat_interleave(f.shape[2]).to(t.float32)
        y = (fields / net.scale.double())
        net.mean = y.mean(dim=0).to(t.float32)
        yc = (y - net.mean.double())
        _, _, vh = t.linalg.svd(yc, full_matrices=False)
        k = min(modes, vh.shape[0])
        P = t.zeros((modes, n), dtype=t.float32, device=fields.device)
        P[:k] = vh[:k].to(t.float32)
        g = t.Generator(device="cpu").manual_seed(seed)
        rnd = lambda *shape: (t.randn(*shape, generator=g) * math.sqrt(1.0 / shape[-1])).to(fields.device)
        net.params = {
            "W1": rnd(width, 2), "b1": t.zeros(width, device=fields.device),
            "W2": rnd(width, width), "b2": t.zeros(width, device=fields.device),
            "W3": rnd(modes, width) * 0.1, "b3": t.zeros(modes, device=fields.device),
            "P": P,
        }
        return net

    def parameter_count(self):
        return sum(int(v.numel()) for v in self.params.values())

    # ---- forward / backward through the MFMA kernel
    def forward(self, x):
        """x: (B, 2) fp32 scaled parameters -> (normalised field (B, n) fp32,
        cache for backward)."""
        t = dev.torch()
        p = self.params
        bf = lambda a: a.to(t.bfloat16)
        xb = bf(x)
        z1 = mm(xb, bf(p["W1"]).t()) + p["b1"]
        h1 = z1 * t.sigmoid(z1)
        z2 = mm(bf(h1), bf(p["W2"]).t()) + p["b2"]
        h2 = z2 * t.sigmoid(z2)
        c = mm(bf(h2), bf(p["W3"]).t()) + p["b3"]
        y = mm(bf(c), bf(p["P"]), ksplit=1) + self.mean
        return y, (xb, z1, h1, z2, h2, c)

    def backward(self, dy, cache):
        """Gradients of the parameters for dL/dy = dy (B x n, fp32)."""
        t = dev.torch()
        p = self.params
        bf = lambda a: a.to(t.bfloat16)
        xb, z1, h1, z2, h2, c = cache
        dyb = bf(dy)
        g = {}
        g["P"] = mm(bf(c).t(), dyb)                                    # K x n
        dc = mm(dyb, bf(p["P"]).t(), ksplit=_ksplit(self.n))             # B x K
        g["b3"] = dc.sum(0)
        g["W3"] = mm(bf(dc).t(), bf(h2))                                 # K x H
        dh2 = mm(bf(dc), bf(p["W3"]))                                    # B x H
        s2 = t.sigmoid(z2)
        dz2 = dh2 * (s2 * (1.0 + z2 * (1.0 - s2)))
        g["b2"] = dz2.sum(0)
        g["W2"] = mm(bf(dz2).t(), bf(h1))
        dh1 = mm(bf(dz2), bf(p["W2"]))
        s1 = t.sigmoid(z1)
        dz1 = dh1 * (s1 * (1.0 + z1 * (1.0 - s1)))
        g["b1"] = dz1.sum(0)
        g["W1"] = mm(bf(dz1).t(), xb)
        return g

    def loss_and_grads(self, x, target):
        """Mean squared error over the normalised field and its gradients."""
        y, cache = self.forward(x)
        diff = y - target
        B = x.shape[0]
        loss = float((diff * diff).mean())
        grads = self.backward(diff * (2.0 / (B * self.n)), cache)
        return loss, grads

    def predict(self, params_phys):
        """FP64 fields (B, n) for physical (Re, M) pairs."""
        t = dev.torch()
        x = t.as_tensor(scale_params(params_phys), dtype=t.float32, device=self.mean.device)
        y, _ = self.forward(x.reshape(-1, 2))
        return y.double() * self.scale.double()


class _MatMul:
    """Autograd product on the matrix cores: C = A B (fp32 out) with bf16
    operands, the backward products dA = dC B^T and dB = A^T dC through the
    same kernel (split-K for the long reductions)."""

    _fn = None

    @classmethod
    def apply(cls, a, b):
        if cls._fn is None:
            t = dev.torch()

            class Fn(t.autograd.Function):
                @staticmethod
                def forward(ctx, a, b):
                    ab, bb = a.to(t.bfloat16), b.to(t.bfloat16)
                    ctx.save_for_backward(ab, bb)
                    return mm(ab, bb, ksplit=_ksplit(ab.shape[1]))

                @staticmethod
                def backward(ctx, g):
                    ab, bb = ctx.saved_tensors
                    gb = g.to(t.bfloat16)
                    da = mm(gb, bb.t(), ksplit=_ksplit(gb.shape[1])) if ctx.needs_input_grad[0] else None
                    db = mm(ab.t(), gb, ksplit=_ksplit(ab.shape[0])) if ctx.needs_input_grad[1] else None
                    return da, db

            cls._fn = Fn
        return cls._fn.apply(a, b)


@dataclass
class ConvGuessNet:
    """The same trunk as GuessNet with a convolutional decoder instead of the
    linear POD-initialised one:
        c (K) -> coarse grid (C0 x jm/16 x im/16) by a dense layer
        4 x [nearest 2x upsampling, 3 x 3 convolution, silu]
        3 x 3 convolution to the 5 variables, + the mean field.
    Padding is periodic along i (the O-mesh's cut) and replicated along j.
    Each convolution is im2col (torch unfold) and one product on the matrix
    cores; backward by autograd through _MatMul."""
    n: int
    im: int
    jm: int
    width: int = 64
    modes: int = 16
    chans: tuple = (32, 32, 16, 16, 16)
    params: dict = field(default_factory=dict)
    scale: object = None
    mean: object = None

    @staticmethod
    def create(fields, im, jm, width=64, modes=16, chans=(32, 32, 16, 16, 16), seed=0):
        t = dev.torch()
        if im % 16 or jm % 16:
            raise ValueError("the convolutional decoder needs im and jm divisible by 16")
        N, n = fields.shape
        net = ConvGuessNet(n=n, im=im, jm=jm, width=width, modes=modes, chans=tuple(chans))
        f = fields.view(N, 5, -1)
        s = f.abs().amax(dim=(0, 2)).clamp_min(1e-12)
        net.scale = s.repeat_interleave(f.shape[2]).to(t.float32)
        net.mean = (fields / net.scale.double()).mean(dim=0).to(t.float32)
        g = t.Generator(device="cpu").manual_seed(seed)
        rnd = lambda *shape, fan: (t.randn(*shape, generator=g) * math.sqrt(1.0 / fan)).to(fields.device)
        c0 = chans[0]
        cells0 = (jm // 16) * (im // 16)
        p = {
            "W1": rnd(width, 2, fan=2), "b1": t.zeros(width, device=fields.device),
            "W2": rnd(width, width, fan=width), "b2": t.zeros(width, device=fields.device),
            "W3": rnd(modes, width, fan=width), "b3": t.zeros(modes, device=fields.device),
            "Wc": rnd(modes, c0 * cells0, fan=modes), "bc": t.zeros(c0 * cells0, device=fields.device),
        }
        for k in range(len(chans) - 1):
            p[f"K{k}"] = rnd(chans[k] * 9, chans[k + 1], fan=chans[k] * 9)
            p[f"k{k}"] = t.zeros(chans[k + 1], device=fields.device)
        p["Kout"] = rnd(chans[-1] * 9, 5, fan=chans[-1] * 9) * 0.1
        p["kout"] = t.zeros(5, device=fields.device)
        for v in p.values():
            v.requires_grad_(True)
        net.params = p
        return net

    def parameter_count(self):
        return sum(int(v.numel()) for v in self.params.values())

    @staticmethod
    def _conv(x, w, b):
        """3 x 3 convolution of x (B, C, H, W): periodic in W, replicated in
        H; w (9 C, Cout)."""
        t = dev.torch()
        F = t.nn.functional
        B, Cc, H, W = x.shape
        xp = F.pad(F.pad(x, (1, 1, 0, 0), mode="circular"), (0, 0, 1, 1), mode="replicate")
        cols = F.unfold(xp, 3)                                     # (B, 9 C, H W)
        cols = cols.transpose(1, 2).reshape(B * H * W, Cc * 9)
        y = _MatMul.apply(cols, w) + b                             # (B H W, Cout)
        return y.view(B, H, W, -1).permute(0, 3, 1, 2)

    def forward_t(self, x):
        """x (B, 2) fp32 -> normalised field (B, n), differentiable."""
        t = dev.torch()
        F = t.nn.functional
        p = self.params
        silu = F.silu
        h = silu(_MatMul.apply(x, p["W1"].t()) + p["b1"])
        h = silu(_MatMul.apply(h, p["W2"].t()) + p["b2"])
        c = _MatMul.apply(h, p["W3"].t()) + p["b3"]
        g = _MatMul.apply(c, p["Wc"]) + p["bc"]
        B = x.shape[0]
        z = g.view(B, self.chans[0], self.jm // 16, self.im // 16)
        for k in range(len(self.chans) - 1):
            z = F.interpolate(z, scale_factor=2, mode="nearest")
            z = silu(self._conv(z, p[f"K{k}"], p[f"k{k}"]))
        z = self._conv(z, p["Kout"], p["kout"])                    # (B, 5, jm, im)
        return z.reshape(B, self.n) + self.mean

    def loss_and_grads(self, x, target):
        for v in self.params.values():
            v.grad = None
        y = self.forward_t(x)
        diff = y - target
        loss = (diff * diff).mean()
        loss.backward()
        return float(loss.detach()), {k: v.grad for k, v in self.params.items()}

    def predict(self, params_phys):
        t = dev.torch()
        x = t.as_tensor(scale_params(params_phys), dtype=t.float32, device=self.mean.device)
        with t.no_grad():
            y = self.forward_t(x.reshape(-1, 2))
        return y.double() * self.scale.double()


def train(net, params_phys, fields, steps=2000, lr=3e-3, log=None):
    """Adam on the FP32 master weights against the normalised training fields
    (fields: N x n FP64 device tensor).  Returns the loss history."""
    t = dev.torch()
    x = t.as_tensor(scale_params(params_phys), dtype=t.float32, device=fields.device)
    target = (fields / net.scale.double()).to(t.float32)
    m = {k: t.zeros_like(v) for k, v in net.params.items()}
    v = {k: t.zeros_like(v) for k, v in net.params.items()}
    b1, b2, eps = 0.9, 0.999, 1e-8
    hist = []
    for it in range(1, steps + 1):
        loss, g = net.loss_and_grads(x, target)
        hist.append(loss)
        a = lr * (0.5 * (1.0 + math.cos(math.pi * it / steps)) * 0.9 + 0.1)   # cosine to lr/10
        with t.no_grad():
            for k, p in net.params.items():
                m[k].mul_(b1).add_(g[k], alpha=1 - b1)
                v[k].mul_(b2).addcmul_(g[k], g[k], value=1 - b2)
                mh = m[k] / (1 - b1 ** it)
                vh = v[k] / (1 - b2 ** it)
                p.sub_(a * mh / (vh.sqrt() + eps))
        if log is not None and (it % max(1, steps // 10) == 0 or it == 1):
            log(f"train step {it}: loss {loss:.3e}")
    return hist


# ------------------------------------------------------------------ baseflows and Newton
def interior_field(flow, w):
    """(n,) view-copy of the interior of a flat state in (variable, j, i)
    order, the layout of the network's output."""
    return w.view(5, flow.nj, flow.ni)[:, 3:-3, 3:-3].reshape(-1).clone()


def set_interior(flow, w, field_):
    w.view(5, flow.nj, flow.ni)[:, 3:-3, 3:-3] = field_.view(5, flow.jm, flow.im)


@dataclass
class Solve:
    seconds: float
    iterations: int
    converged: bool
    start: str
    norms: list


def solve(flow, w, tol=1e-10, ite=40, cfl=1e21):
    """Newton to max_e ||R_e||_2 < tol from the state in w (in place)."""
    from .newton import newton

    t = dev.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    log = newton(flow, w, ite=ite, cfl=cfl, tol=tol)
    t.cuda.synchronize()
    return Solve(time.perf_counter() - t0, len(log.norms) - 1, log.converged, "", log.norms)


def make_flow(im, jm, re, mach):
    from .cases import cylinder_case
    from .flow import Flow

    return Flow(cylinder_case(im, jm, mach=float(mach), reynolds=float(re), noise=0.0))


def freestream(flow):
    from .flow import to_device

    return to_device(flow.case.w)


def continuation_solve(im, jm, re, mach, tol=1e-10, ite=12, re_start=RE_RANGE[0], log=None):
    """Continuation Newton in Re (broadcast_amd.newton.continuation) from the
    free stream at re_start to re at fixed M: the fallback when Newton from
    a start does not converge.  Returns (Solve, flow, w)."""
    from .newton import continuation

    t = dev.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    f, w, hist = continuation(lambda r: make_flow(im, jm, r, mach), float(re), float(re_start), steps=4, ite=ite,
                              cfl=1e21, tol=tol)
    t.cuda.synchronize()
    its = sum(len(lg.norms) - 1 for _, lg in hist)
    if log is not None:
        log(f"continuation Re {re_start:.0f} -> {re:.1f}: {len(hist)} solves, {its} Newton steps")
    return Solve(time.perf_counter() - t0, its, True, "continuation", hist[-1][1].norms), f, w


def baseflows(im, jm, params_phys, tol=1e-10, log=None, warm=True, ite=12):
    """Newton baseflows at each (Re, M): (departures from the free stream,
    N x n FP64 device tensor, list of Solve, the (Re, M) pairs).  With warm,
    the points are taken in order of Re and each starts from the converged
    baseflow nearest in scaled (Re, M) (the free stream for the first); a
    start that does not converge in `ite` steps falls back to continuation
    in Re.  A pair that converges in neither way is left out and logged."""
    t = dev.torch()
    out, stats, used = [], [], []
    order = np.argsort(np.asarray(params_phys)[:, 0]) if warm else np.arange(len(params_phys))
    for k in order:
        re, ma = params_phys[k]
        f = make_flow(im, jm, re, ma)
        w = freestream(f)
        free = interior_field(f, w)
        start = "free stream"
        if warm and used:
            d = np.linalg.norm(scale_params(used) - scale_params([[re, ma]]), axis=1)
            near = int(np.argmin(d))
            set_interior(f, w, free + out[near])
            start = f"baseflow at Re {used[near][0]:.1f} M {used[near][1]:.3f}"
        s = solve(f, w, tol=tol, ite=ite)
        s.start = start
        if not s.converged:
            try:
                s, f, w = continuation_solve(im, jm, re, ma, tol=tol, ite=ite, log=log)
            except RuntimeError as err:
                if log is not None:
                    log(f"baseflow Re {re:.1f} M {ma:.3f}: {err}, left out")
                continue
        if log is not None:
            log(f"baseflow {len(used)}: Re {re:.1f} M {ma:.3f} from {s.start}: {s.iterations} Newton steps "
                f"{s.seconds:.2f} s")
        out.append(interior_field(f, w) - free)
        stats.append(s)
        used.append((re, ma))
    if not out:
        raise RuntimeError("no baseflow converged")
    return t.stack(out), stats, np.array(used)


def guess_field(flow, net, re, mach):
    """Free stream of (Re, M) + the network's departure, (n,) FP64."""
    return interior_field(flow, freestream(flow)) + net.predict([[re, mach]])[0]


def solve_from_guess(flow, guess, tol=1e-10, ite=12):
    """Newton from the learned guess (interior set, ghosts by the BCs);
    continuation Newton in Re is the fallback when that does not converge.
    The time includes the fallback's."""
    t = dev.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    w = freestream(flow)
    set_interior(flow, w, guess)
    s = solve(flow, w, tol=tol, ite=ite)
    s.start = "learned"
    if not s.converged or not all(np.isfinite(s.norms)):
        ph = flow.case.phys
        s2, _, w = continuation_solve(flow.im, flow.jm, ph.reynolds, ph.mach, tol=tol, ite=ite)
        s = Solve(time.perf_counter() - t0, s.iterations + s2.iterations, s2.converged, "learned, fell back",
                  s.norms + s2.norms)
    else:
        s.seconds = time.perf_counter() - t0
    return s, w


def solve_plain(flow, tol=1e-10, ite=12):
    """The plain method: Newton from the free stream, continuation Newton in
    Re when that does not converge in `ite` steps (time included)."""
    t = dev.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    w = freestream(flow)
    s = solve(flow, w, tol=tol, ite=ite)
    s.start = "free stream"
    if not s.converged or not all(np.isfinite(s.norms)):
        ph = flow.case.phys
        s2, _, w = continuation_solve(flow.im, flow.jm, ph.reynolds, ph.mach, tol=tol, ite=ite)
        s = Solve(time.perf_counter() - t0, s.iterations + s2.iterations, s2.converged, "free stream, fell back",
                  s.norms + s2.norms)
    else:
        s.seconds = time.perf_counter() - t0
    return s, w


# ------------------------------------------------------------------ experiment
def experiment(im, jm, n_train=96, n_test=16, width=64, modes=16, steps=3000, seed=4, log=print, progress=None):
    """Baseflows at n_train Sobol points, the network trained on them, then
    for n_test held-out Sobol points Newton to ||R|| < 1e-10 from the free
    stream and from the learned guess.  Returns a JSON-able summary;
    progress(rows), when given, receives the held-out rows after each one."""
    t = dev.torch()
    params = sobol_params(n_train + n_test, seed=seed)
    train_p, test_p = params[:n_train], params[n_train:]
    t0 = time.perf_counter()
    fields, stats, train_p = baseflows(im, jm, train_p, log=log)
    gen_s = time.perf_counter() - t0
    net = GuessNet.create(fields, width=width, modes=modes, seed=seed)
    t.cuda.synchronize()
    t0 = time.perf_counter()
    hist = train(net, train_p, fields, steps=steps, log=log)
    t.cuda.synchronize()
    train_s = time.perf_counter() - t0
    rows = []
    for re, ma in test_p:
        f = make_flow(im, jm, re, ma)
        free = interior_field(f, freestream(f))
        plain, w0 = solve_plain(f)
        t.cuda.synchronize()
        t1 = time.perf_counter()
        guess = guess_field(f, net, re, ma)       # inference is part of the learned start's time
        t.cuda.synchronize()
        infer_s = time.perf_counter() - t1
        learned, wl = solve_from_guess(f, guess)
        sol = interior_field(f, wl)
        rows.append({
            "re": float(re), "mach": float(ma),
            "plain": {"seconds": plain.seconds, "newton_steps": plain.iterations, "converged": plain.converged,
                      "start": plain.start},
            "learned": {"seconds": learned.seconds + infer_s, "newton_steps": learned.iterations,
                        "converged": learned.converged, "start": learned.start, "inference_s": infer_s},
            "distance_to_baseflow": {"free_stream": float((free - sol).norm() / sol.norm()),
                                     "guess": float((guess - sol).norm() / sol.norm())},
            "same_baseflow": float((interior_field(f, w0) - sol).norm() / sol.norm()),
        })
        log(f"held out Re {re:.1f} M {ma:.3f}: plain {plain.iterations} steps {plain.seconds:.2f} s "
            f"({plain.start}), learned {learned.iterations} steps {learned.seconds + infer_s:.2f} s "
            f"({learned.start}), guess distance {rows[-1]['distance_to_baseflow']['guess']:.2e}")
        import json as _json
        log("ROW " + _json.dumps(rows[-1]))
        if progress is not None:
            progress(rows)
    ps = np.array([r["plain"]["seconds"] for r in rows])
    ls = np.array([r["learned"]["seconds"] for r in rows])
    pi = np.array([r["plain"]["newton_steps"] for r in rows])
    li = np.array([r["learned"]["newton_steps"] for r in rows])
    # headline: the points where both starts reached the tolerance; a plain
    # solve stopped at its iteration cap would count at that capped time
    both = np.array([r["plain"]["converged"] and r["learned"]["converged"] for r in rows], dtype=bool)
    unconv = [{"re": r["re"], "mach": r["mach"], "plain_converged": r["plain"]["converged"],
               "learned_converged": r["learned"]["converged"]} for r, b in zip(rows, both) if not b]
    return {
        "mesh": [im, jm], "train": int(len(train_p)), "train_requested": n_train, "held_out": n_test,
        "sobol_seed": seed,
        "network": {"width": width, "modes": modes, "parameters": net.parameter_count(),
                    "dtype": "bf16 operands on MFMA, fp32 accumulation, fp32 master weights"},
        "baseflow_generation_s": gen_s, "training_s": train_s, "training_steps": steps,
        "loss_first_last": [hist[0], hist[-1]],
        "training_newton_steps_mean": float(np.mean([s.iterations for s in stats])),
        "training_starts": "in order of Re, each from the nearest converged baseflow (continuation in Re as "
                           "fallback)",
        "held_out_results": rows,
        "speedup_wall_converged": float(ps[both].sum() / ls[both].sum()) if both.any() else None,
        "converged_points": int(both.sum()), "not_converged": unconv,
        "plain_seconds_converged": float(ps[both].sum()), "learned_seconds_converged": float(ls[both].sum()),
        "speedup_wall_median_converged": float(np.median(ps[both] / ls[both])) if both.any() else None,
        "plain_seconds_total": float(ps.sum()), "learned_seconds_total": float(ls.sum()),
        "speedup_wall_total_incl_capped": float(ps.sum() / ls.sum()),
        "newton_steps_mean_plain": float(pi[both].mean()) if both.any() else None,
        "newton_steps_mean_learned": float(li[both].mean()) if both.any() else None,
        "all_converged": bool(all(r["plain"]["converged"] and r["learned"]["converged"] for r in rows)),
        "fallbacks": int(sum(r["learned"]["start"] != "learned" for r in rows)),
    }


def compare_decoders(im, jm, n_train=96, n_test=16, width=64, modes=16, steps=3000, seed=4, log=print,
                     progress=None):
    """The linear POD-initialised decoder (GuessNet) against the
    convolutional one (ConvGuessNet), both trained on the same baseflows with
    the same trunk, steps and schedule; for each held-out point Newton from
    the free stream and from each guess.  Returns a JSON-able summary."""
    t = dev.torch()
    params = sobol_params(n_train + n_test, seed=seed)
    train_p, test_p = params[:n_train], params[n_train:]
    t0 = time.perf_counter()
    fields, stats, train_p = baseflows(im, jm, train_p, log=log)
    gen_s = time.perf_counter() - t0
    nets, info = {}, {}
    for name in ("linear", "conv"):
        if name == "linear":
            net = GuessNet.create(fields, width=width, modes=modes, seed=seed)
        else:
            net = ConvGuessNet.create(fields, im, jm, width=width, modes=modes, seed=seed)
        t.cuda.synchronize()
        t0 = time.perf_counter()
        hist = train(net, train_p, fields, steps=steps, log=log)
        t.cuda.synchronize()
        nets[name] = net
        info[name] = {"parameters": net.parameter_count(), "training_s": time.perf_counter() - t0,
                      "loss_first_last": [hist[0], hist[-1]]}
        log(f"{name} decoder: {net.parameter_count()} parameters, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    rows = []
    for re, ma in test_p:
        f = make_flow(im, jm, re, ma)
        plain, w0 = solve_plain(f)
        sol = interior_field(f, w0)
        row = {"re": float(re), "mach": float(ma),
               "plain": {"seconds": plain.seconds, "newton_steps": plain.iterations, "converged": plain.converged}}
        for name, net in nets.items():
            t.cuda.synchronize()
            t1 = time.perf_counter()
            guess = guess_field(f, net, re, ma)
            t.cuda.synchronize()
            infer_s = time.perf_counter() - t1
            dist = float((guess - sol).norm() / sol.norm())
            s, _ = solve_from_guess(f, guess)
            row[name] = {"seconds": s.seconds + infer_s, "newton_steps": s.iterations, "converged": s.converged,
                         "start": s.start, "guess_distance": dist}
        rows.append(row)
        log(f"held out Re {re:.1f} M {ma:.3f}: plain {plain.iterations} steps {plain.seconds:.2f} s; "
            + "; ".join(f"{n} {row[n]['newton_steps']} steps {row[n]['seconds']:.2f} s "
                        f"(distance {row[n]['guess_distance']:.2e})" for n in nets))
        if progress is not None:
            progress(rows)
    out = {"mesh": [im, jm], "train": int(len(train_p)), "held_out": n_test, "sobol_seed": seed,
           "trunk": {"width": width, "modes": modes}, "training_steps": steps, "baseflow_generation_s": gen_s,
           "held_out_results": rows}
    ps = np.array([r["plain"]["seconds"] for r in rows])
    for name in nets:
        ok = np.array([r["plain"]["converged"] and r[name]["converged"] for r in rows], dtype=bool)
        ls = np.array([r[name]["seconds"] for r in rows])
        out[name] = dict(info[name], **{
            "guess_distance_median": float(np.median([r[name]["guess_distance"] for r in rows])),
            "guess_distance_max": float(np.max([r[name]["guess_distance"] for r in rows])),
            "newton_steps_mean": float(np.mean([r[name]["newton_steps"] for r in rows])),
            "seconds_total": float(ls.sum()),
            "speedup_wall_converged": float(ps[ok].sum() / ls[ok].sum()) if ok.any() else None,
            "converged_points": int(ok.sum()),
            "fallbacks": int(sum(r[name]["start"] != "learned" for r in rows)),
        })
    out["plain"] = {"newton_steps_mean": float(np.mean([r["plain"]["newton_steps"] for r in rows])),
                    "seconds_total": float(ps.sum())}
    return out


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="learned Newton initial guess on the cylinder O-mesh")
    ap.add_argument("--im", type=int, default=192)
    ap.add_argument("--jm", type=int, default=128)
    ap.add_argument("--train", type=int, default=96)
    ap.add_argument("--test", type=int, default=16)
    ap.add_argument("--width", type=int, default=64)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--compare-decoders", action="store_true",
                    help="train the linear and the convolutional decoder on the same baseflows and compare them")
    a = ap.parse_args(argv)
    def progress(rows):
        # the held-out rows so far, so that a run cut short keeps them
        if a.out:
            with open(a.out, "w") as fh:
                json.dump({"partial": True, "mesh": [a.im, a.jm], "held_out_results": rows}, fh, indent=1)

    run = compare_decoders if a.compare_decoders else experiment
    res = run(a.im, a.jm, a.train, a.test, a.width, a.modes, a.steps, log=lambda s: print(s, flush=True),
              progress=progress)
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(json.dumps({k: v for k, v in res.items() if k != "held_out_results"}), flush=True)


if __name__ == "__main__":
    main()
