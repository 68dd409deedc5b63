"""Learned SIRT: a small BF16 CNN correction interleaved with SIRT steps.

No reference counterpart.  The correction network

    y = x + conv4(relu(conv3(relu(conv2(relu(conv1(x)))))))

(3x3 convolutions, 1 -> 32 -> 32 -> 32 -> 1 channels, 19,105 parameters) runs entirely on
the hand-written kernels of sk_learn.hip: the 32 -> 32 layers as MFMA implicit GEMMs, the thin
first/last layers on VALU, BF16 activations with FP32 accumulation; backward is the same
kernels (data gradients) plus deterministic weight-gradient reductions.  torch holds the
FP32 master weights, their optimizer and the autograd graph.

``LearnedSIRT`` alternates one SIRT step and one correction ``stages`` times; if a
correction produces a non-finite image it is dropped for that step, and if the final data
residual ||Ax - b|| / ||b|| exceeds plain SIRT's after the same number of steps the plain
SIRT result is returned instead.
"""
from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np

import torch

from . import _native as nat
from .autograd import SirtWeights, _SirtStep
from .projector import Context, Geometry, _ptr, _stream

NCH = 32
BF16 = torch.bfloat16


def _chk(rc, ctx, what):
    nat.check(rc, ctx.handle, what)


_BF16_CACHE: dict = {}


def _bf16_of(w: torch.Tensor) -> torch.Tensor:
    """w in BF16, contiguous; cached per tensor (a weak reference: a new tensor at a recycled
    id is never mistaken for it) until its data or its version counter changes."""
    key = id(w)
    hit = _BF16_CACHE.get(key)
    stamp = (w.data_ptr(), w._version, tuple(w.shape))
    if hit is None or hit[0]() is not w or hit[1] != stamp:
        hit = (weakref.ref(w), stamp, w.detach().to(BF16).contiguous())
        _BF16_CACHE[key] = hit
    return hit[2]


class _Correction(torch.autograd.Function):
    """y = res + conv4(relu(conv3(relu(conv2(relu(conv1(inp))))))); inp is the image [B, H, W]
    (one input channel) or the image and its SIRT update direction [B, H, W, 2]."""

    @staticmethod
    def forward(fctx, inp, res, w1, b1, w2, b2, w3, b3, w4, b4, ctx: Context):
        L = nat.lib()
        inp = inp.contiguous()
        res = res.contiguous()
        B, H, W = res.shape
        cin = 1 if inp.dim() == 3 else inp.shape[-1]
        dev = res.device
        s = _stream(ctx.device)
        # BF16 copies of the weights: w1 [32, 3, 3] or [32, 3, 3, 2], w2 and w3 [32, 3, 3, 32],
        # w4 [3, 3, 32]; converted again only when a weight has changed
        w1b, w2b, w3b, w4b = (_bf16_of(w) for w in (w1, w2, w3, w4))
        h1 = torch.empty((B, H, W, NCH), dtype=BF16, device=dev)
        h2 = torch.empty_like(h1)
        h3 = torch.empty_like(h1)
        y = torch.empty_like(res)
        h = ctx.handle
        conv_in = L.sk_conv_in if cin == 1 else L.sk_conv_in2
        _chk(conv_in(h, B, H, W, _ptr(inp), _ptr(w1b), _ptr(b1), _ptr(h1), s), ctx, "conv1")
        _chk(L.sk_conv32(h, B, H, W, _ptr(h1), None, _ptr(w2b), _ptr(b2), _ptr(h2), s), ctx, "conv2")
        _chk(L.sk_conv32(h, B, H, W, _ptr(h2), None, _ptr(w3b), _ptr(b3), _ptr(h3), s), ctx, "conv3")
        _chk(L.sk_conv_out(h, B, H, W, _ptr(h3), _ptr(w4b), _ptr(b4), _ptr(res), _ptr(y), s), ctx, "conv4")
        fctx.ctx = ctx
        fctx.cin = cin
        fctx.save_for_backward(inp, h1, h2, h3, w1b, w2b, w3b, w4b)
        return y

    @staticmethod
    def backward(fctx, gy):
        ctx = fctx.ctx
        cin = fctx.cin
        L = nat.lib()
        h = ctx.handle
        s = _stream(ctx.device)
        inp, h1, h2, h3, w1b, w2b, w3b, w4b = fctx.saved_tensors
        gy = gy.contiguous().float()
        B, H, W = gy.shape
        dev = gy.device

        def wgrad(kind, g, mask, x, n):
            out = torch.empty(n, dtype=torch.float32, device=dev)
            _chk(L.sk_conv_wgrad(h, kind, B, H, W, _ptr(g), None if mask is None else _ptr(mask),
                                 _ptr(x), _ptr(out), s), ctx, "wgrad")
            return out

        # conv4 (32 -> 1, residual)
        g4 = wgrad(2, gy, None, h3, 9 * NCH + 1)
        dh3 = torch.empty_like(h3)
        _chk(L.sk_conv_out_dgrad(h, B, H, W, _ptr(gy), _ptr(w4b), _ptr(dh3), s), ctx, "conv4 dgrad")
        # conv3, conv2 (32 -> 32): gradient wrt the pre-activation = dh * (h > 0)
        g3 = wgrad(1, dh3, h3, h2, NCH * 9 * NCH + NCH)
        w3t = w3b.flip(1, 2).permute(3, 1, 2, 0).contiguous()
        dh2 = torch.empty_like(h2)
        _chk(L.sk_conv32(h, B, H, W, _ptr(dh3), _ptr(h3), _ptr(w3t), None, _ptr(dh2), s), ctx, "conv3 dgrad")
        g2 = wgrad(1, dh2, h2, h1, NCH * 9 * NCH + NCH)
        w2t = w2b.flip(1, 2).permute(3, 1, 2, 0).contiguous()
        dh1 = torch.empty_like(h1)
        _chk(L.sk_conv32(h, B, H, W, _ptr(dh2), _ptr(h2), _ptr(w2t), None, _ptr(dh1), s), ctx, "conv2 dgrad")
        # conv1 (cin -> 32)
        n1 = NCH * 9 * cin
        g1 = wgrad(0 if cin == 1 else 3, dh1, h1, inp, n1 + NCH)
        dinp = torch.empty_like(inp)
        dgrad = L.sk_conv_in_dgrad if cin == 1 else L.sk_conv_in2_dgrad
        _chk(dgrad(h, B, H, W, _ptr(dh1), _ptr(h1), _ptr(w1b), _ptr(dinp), s), ctx, "conv1 dgrad")
        n2, n4 = NCH * 9 * NCH, 9 * NCH
        gw1 = g1[:n1].view(NCH, 3, 3) if cin == 1 else g1[:n1].view(NCH, 3, 3, cin)
        return (dinp, gy, gw1, g1[n1:], g2[:n2].view(NCH, 3, 3, NCH), g2[n2:],
                g3[:n2].view(NCH, 3, 3, NCH), g3[n2:], g4[:n4].view(3, 3, NCH), g4[n4:], None)


class Correction(torch.nn.Module):
    """The correction network; call on FP32 images [B, H, W] on the context's device.  It also
    holds the per-stage relaxation factors of the SIRT steps it is interleaved with."""

    MAX_STAGES = 16

    def __init__(self, ctx: Context | None = None, seed: int = 0, cin: int = 1):
        super().__init__()
        if cin not in (1, 2):
            raise ValueError("cin must be 1 (image) or 2 (image and SIRT update direction)")
        self.ctx = ctx or Context.default(0)
        self.cin = cin
        g = torch.Generator().manual_seed(seed)

        def kaiming(shape, fan_in, gain=math.sqrt(2.0)):
            return torch.randn(shape, generator=g) * (gain / math.sqrt(fan_in))

        dev = torch.device("cuda", self.ctx.device)
        self.w1 = torch.nn.Parameter(kaiming((NCH, 3, 3) if cin == 1 else (NCH, 3, 3, cin), 9 * cin).to(dev))
        self.b1 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        self.w2 = torch.nn.Parameter(kaiming((NCH, 3, 3, NCH), 9 * NCH).to(dev))
        self.b2 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        self.w3 = torch.nn.Parameter(kaiming((NCH, 3, 3, NCH), 9 * NCH).to(dev))
        self.b3 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        # the last layer starts at zero: the network begins as the identity
        self.w4 = torch.nn.Parameter(torch.zeros((3, 3, NCH), device=dev))
        self.b4 = torch.nn.Parameter(torch.zeros(1, device=dev))
        # relaxation of the SIRT step of each stage k >= 1 (x + lam_k * (step(x) - x)); stage 0
        # takes the plain step, which the fallback rule's plain SIRT shares
        self.relax = torch.nn.Parameter(torch.ones(self.MAX_STAGES, device=dev))

    def forward(self, x: torch.Tensor, d: torch.Tensor | None = None) -> torch.Tensor:
        """x: images [B, H, W] (or [H, W]); d: their SIRT update directions (two-channel
        networks only)."""
        squeeze = x.dim() == 2
        xb = x.unsqueeze(0) if squeeze else x
        if self.cin == 2:
            if d is None:
                raise ValueError("a two-channel correction needs the SIRT update direction")
            db = d.unsqueeze(0) if squeeze else d
            inp = torch.stack([xb, db], dim=-1)
        else:
            inp = xb
        y = _Correction.apply(inp, xb, self.w1, self.b1, self.w2, self.b2, self.w3, self.b3, self.w4,
                              self.b4, self.ctx)
        return y.squeeze(0) if squeeze else y

    def save(self, path: str, **meta) -> None:
        arrs = {k: v.detach().cpu().numpy() for k, v in self.state_dict().items()}
        arrs.update({f"meta_{k}": np.asarray(v) for k, v in meta.items()})
        np.savez(path, **arrs)

    def load(self, path: str) -> dict:
        z = np.load(path, allow_pickle=False)
        sd = {k: torch.from_numpy(z[k]) for k in z.files if not k.startswith("meta_")}
        if "relax" not in sd:  # weights trained without stage relaxation: plain steps
            sd["relax"] = torch.ones(self.MAX_STAGES)
        cin = 1 if sd["w1"].dim() == 3 else int(sd["w1"].shape[-1])
        if cin != self.cin:  # the file decides the input channels
            self.cin = cin
            self.w1 = torch.nn.Parameter(torch.empty(tuple(sd["w1"].shape), device=self.w1.device))
        self.load_state_dict(sd)
        return {k[5:]: z[k] for k in z.files if k.startswith("meta_")}


# ---------------------------------------------------------------------------- data
def gpu_phantoms(ctx: Context, rng: np.random.Generator, count: int, N: int, n_ellipses: int = 10,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """``count`` phantoms [count, N, N] rendered on the GPU; draws follow
    synkrotomo_amd.phantom.ellipse_params, so the images equal phantom.render's."""
    from .phantom import ellipse_params
    params = np.empty((count, n_ellipses, 7), np.float64)
    for i in range(count):
        c, a, r, v = ellipse_params(rng, n_ellipses)
        params[i, :, 0:2] = c
        params[i, :, 2:4] = a
        params[i, :, 4] = np.cos(r)
        params[i, :, 5] = np.sin(r)
        params[i, :, 6] = v
    dev = torch.device("cuda", ctx.device)
    p = torch.from_numpy(params).to(dev)
    if out is None:
        out = torch.empty((count, N, N), dtype=torch.float32, device=dev)
    nat.check(nat.lib().sk_phantoms(ctx.handle, count, N, n_ellipses, _ptr(p), _ptr(out), _stream(ctx.device)),
              ctx.handle, "phantoms")
    return out


def noisy_sinogram(clean: torch.Tensor, scale: float, I0: float = 1e4,
                   gen: torch.Generator | None = None) -> torch.Tensor:
    """Poisson transmission noise: counts ~ Poisson(I0 exp(-scale * Ax)),
    b = -log(max(counts, 1) / I0) / scale."""
    lam = I0 * torch.exp(-scale * clean)
    counts = torch.poisson(lam, generator=gen)
    return -torch.log(torch.clamp(counts, min=1.0) / I0) / scale


def residual(geom: Geometry, x: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Per-image relative data residual ||A x - b|| / ||b||."""
    return _residual_of(geom.fp(x.contiguous()), b)


def _residual_of(ax: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    B = b.shape[0] if b.dim() == 3 else 1
    d = (ax - b).reshape(B, -1).double()
    return d.norm(dim=1) / b.reshape(B, -1).double().norm(dim=1)


# ---------------------------------------------------------------------------- model
class LearnedSIRT:
    """``stages`` x (SIRT step, correction) with the plain-SIRT fallback.

    Stage k takes the SIRT step relaxed by the network's factor lam_k (k >= 1; the first step
    is plain), then the correction.  ``tail_`` continues a reconstruction with SIRT steps
    relaxed by ``tail_relax`` (1 = the reference's SIRT)."""

    def __init__(self, geom: Geometry, net: Correction, stages: int = 5, tail_relax: float = 1.0):
        if not 1 <= stages <= Correction.MAX_STAGES:
            raise ValueError(f"stages must be in [1, {Correction.MAX_STAGES}]")
        if not 0.0 < tail_relax < 2.0:
            raise ValueError("tail_relax must be in (0, 2)")
        self.geom, self.net, self.stages, self.tail_relax = geom, net, stages, float(tail_relax)
        self.w = SirtWeights(geom)

    def sirt_step(self, x, b):
        return _SirtStep.apply(x, b, self.geom, self.w.rowinv, self.w.colinv)

    def stage_step(self, x, b, k: int):
        """The SIRT step of stage k: x + lam_k (step(x) - x), lam_0 = 1."""
        s = self.sirt_step(x, b)
        return s if k == 0 else x + self.net.relax[k] * (s - x)

    def stage(self, x, b, k: int):
        """Stage k: the relaxed SIRT step, then the correction (which also sees the step's
        update direction when it has two input channels)."""
        xs = self.stage_step(x, b, k)
        return self.net(xs, xs - x) if self.net.cin == 2 else self.net(xs)

    @torch.no_grad()
    def tail_(self, x: torch.Tensor, b: torch.Tensor, iters: int, relax: float | None = None) -> torch.Tensor:
        """``iters`` SIRT steps in place on x [B, N, N], each relaxed by ``relax`` (default:
        tail_relax): x + lam C B(R (b - A x)), with lam folded into C."""
        lam = self.tail_relax if relax is None else float(relax)
        if iters <= 0:
            return x
        if lam == 1.0:
            return self.geom.sirt_(x, b, iters)
        key = (lam, self.w.colinv.data_ptr())
        if getattr(self, "_tail_key", None) != key:
            self._tail_colinv, self._tail_key = self.w.colinv * lam, key
        q = torch.empty(b.shape, dtype=torch.float32, device=b.device)
        for _ in range(iters):
            self.geom.sirt_step_(x, b, self.w.rowinv, self._tail_colinv, q)
        return x

    def unrolled(self, b: torch.Tensor, checkpoint: bool = True, all_stages: bool = False):
        """Differentiable forward used for training: images [B, N, N] after the last stage
        (or the list of every stage's output)."""
        from torch.utils.checkpoint import checkpoint as ck
        B = b.shape[0]
        x = torch.zeros((B, self.geom.N, self.geom.N), device=b.device)
        outs = []
        for k in range(self.stages):
            if checkpoint and torch.is_grad_enabled():
                x = ck(lambda xx, bb, kk=k: self.stage(xx, bb, kk), x, b, use_reentrant=False)
            else:
                x = self.stage(x, b, k)
            outs.append(x)
        return outs if all_stages else x

    @torch.no_grad()
    def reconstruct(self, b: torch.Tensor, guard: bool = True):
        """The stages, each correction that yields a non-finite image dropped for its stage.
        guard: also run ``stages`` plain SIRT steps and return them for every image whose data
        residual they beat (the equal-cost fallback rule).  Returns (images, fell_back [B])."""
        B = b.shape[0]
        x = torch.zeros((B, self.geom.N, self.geom.N), device=b.device)
        xp = None
        b2 = torch.cat((b, b)) if guard and self.stages > 1 else None
        for k in range(self.stages):
            if guard and k > 0:
                # the stage's SIRT step and plain SIRT's next step in one call of 2B images
                # (each image's step is independent of the others in the batch: the same bits
                # as two calls, in one launch pair that fills the chip better)
                both = torch.cat((x, xp))
                self.geom.sirt_step_(both, b2, self.w.rowinv, self.w.colinv)
                s, xp = both[:B], both[B:]
                xs = x + self.net.relax[k] * (s - x)  # stage_step's relaxation
            else:
                xs = self.stage_step(x, b, k)
                if guard:
                    # both paths start from zero, so plain SIRT's first step is the first
                    # stage's (the kernels are deterministic: the same bits)
                    xp = xs.clone()
            xc = self.net(xs, xs - x) if self.net.cin == 2 else self.net(xs)
            # a non-finite pixel makes its image's sum non-finite (a float sum reduction, cheaper
            # than reducing a per-pixel test); a finite image whose sum overflows is dropped too
            bad = ~torch.isfinite(xc.reshape(B, -1).sum(dim=1))
            x = torch.where(bad.view(B, 1, 1), xs, xc)
        if not guard:
            return x, torch.zeros(B, dtype=torch.bool, device=b.device)
        # both residuals from one projection call of 2B images (per-image results unchanged)
        ax = self.geom.fp(torch.cat((x, xp)))
        worse = _residual_of(ax[:B], b) > _residual_of(ax[B:], b)
        return torch.where(worse.view(B, 1, 1).to(x.device), xp, x), worse

    @torch.no_grad()
    def solve(self, b: torch.Tensor, target: float | torch.Tensor, max_iters: int,
              check_every: int = 1):
        """Reconstruct to a data residual: the stages, then SIRT steps relaxed by tail_relax
        until ||A x - b|| / ||b|| <= target (a number or one per image), each image kept at the
        first check where it is at or below its target.  The residual judges the output: an
        image that has not reached its target after ``max_iters`` iterations in all (the stages
        count as one each) falls back to plain SIRT from zero with the same budget, and the
        result with the lower residual is returned.
        Returns (images [B, N, N], iterations [B] int64, fell_back [B] bool)."""
        if max_iters < self.stages:
            raise ValueError("max_iters must cover the stages")
        B = b.shape[0]
        tgt = torch.as_tensor(target, dtype=torch.float64, device=b.device).expand(B)
        x, _ = self.reconstruct(b, guard=False)
        out = x.clone()
        iters = torch.full((B,), -1, dtype=torch.int64, device=b.device)
        k = self.stages
        while True:
            hit = (iters < 0) & (residual(self.geom, x, b) <= tgt)
            out[hit] = x[hit]
            iters[hit] = k
            if bool((iters >= 0).all()) or k >= max_iters:
                break
            n = min(check_every, max_iters - k)
            self.tail_(x, b, n)
            k += n
        fell = iters < 0
        if bool(fell.any()):
            idx = fell.nonzero().flatten()
            xp = torch.zeros((idx.numel(), self.geom.N, self.geom.N), device=b.device)
            bp = b[idx].contiguous()
            self.geom.sirt_(xp, bp, max_iters)
            better = residual(self.geom, xp, bp) < residual(self.geom, x[idx], bp)
            out[idx] = torch.where(better.view(-1, 1, 1), xp, x[idx])
            iters[idx] = max_iters
            fell[idx] = better
        return out, iters, fell
