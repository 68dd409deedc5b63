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

import numpy as np
import torch

from . import _native as nat
from .autograd import SirtWeights, _SirtStep
from .projector import Context, Geometry, _ptr, _stream

NCH = 32
BF16 = torch.bfloat16


def _chk(rc, ctx, what):
    nat.check(rc, ctx.handle, what)


class _Correction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, w1, b1, w2, b2, w3, b3, w4, b4, ctx: Context):
        L = nat.lib()
        x = x.contiguous()
        B, H, W = x.shape
        dev = x.device
        s = _stream(ctx.device)
        w1b = w1.to(BF16).contiguous()          # [32, 3, 3]
        w2b = w2.to(BF16).contiguous()          # [32, 3, 3, 32]
        w3b = w3.to(BF16).contiguous()
        w4b = w4.to(BF16).contiguous()          # [3, 3, 32]
        h1 = torch.empty((B, H, W, NCH), dtype=BF16, device=dev)
        h2 = torch.empty_like(h1)
        h3 = torch.empty_like(h1)
        y = torch.empty_like(x)
        h = ctx.handle
        _chk(L.sk_conv_in(h, B, H, W, _ptr(x), _ptr(w1b), _ptr(b1), _ptr(h1), s), ctx, "conv1")
        _chk(L.sk_conv32(h, B, H, W, _ptr(h1), None, _ptr(w2b), _ptr(b2), _ptr(h2), s), ctx, "conv2")
        _chk(L.sk_conv32(h, B, H, W, _ptr(h2), None, _ptr(w3b), _ptr(b3), _ptr(h3), s), ctx, "conv3")
        _chk(L.sk_conv_out(h, B, H, W, _ptr(h3), _ptr(w4b), _ptr(b4), _ptr(x), _ptr(y), s), ctx, "conv4")
        fctx.ctx = ctx
        fctx.save_for_backward(x, h1, h2, h3, w1b, w2b, w3b, w4b)
        return y

    @staticmethod
    def backward(fctx, gy):
        ctx = fctx.ctx
        L = nat.lib()
        h = ctx.handle
        s = _stream(ctx.device)
        x, h1, h2, h3, w1b, w2b, w3b, w4b = fctx.saved_tensors
        gy = gy.contiguous().float()
        B, H, W = gy.shape
        dev = gy.device

        def wgrad(kind, g, mask, inp, n):
            out = torch.empty(n, dtype=torch.float32, device=dev)
            _chk(L.sk_conv_wgrad(h, kind, B, H, W, _ptr(g), None if mask is None else _ptr(mask),
                                 _ptr(inp), _ptr(out), s), ctx, "wgrad")
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
        # conv1 (1 -> 32)
        g1 = wgrad(0, dh1, h1, x, NCH * 9 + NCH)
        dx = torch.empty_like(x)
        _chk(L.sk_conv_in_dgrad(h, B, H, W, _ptr(dh1), _ptr(h1), _ptr(w1b), _ptr(dx), s), ctx, "conv1 dgrad")
        dx = dx + gy  # residual path
        n1, n2, n4 = NCH * 9, NCH * 9 * NCH, 9 * NCH
        return (dx, g1[:n1].view(NCH, 3, 3), g1[n1:], g2[:n2].view(NCH, 3, 3, NCH), g2[n2:],
                g3[:n2].view(NCH, 3, 3, NCH), g3[n2:], g4[:n4].view(3, 3, NCH), g4[n4:], None)


class Correction(torch.nn.Module):
    """The correction network; call on FP32 images [B, H, W] on the context's device."""

    def __init__(self, ctx: Context | None = None, seed: int = 0):
        super().__init__()
        self.ctx = ctx or Context.default(0)
        g = torch.Generator().manual_seed(seed)

        def kaiming(shape, fan_in, gain=math.sqrt(2.0)):
            return torch.randn(shape, generator=g) * (gain / math.sqrt(fan_in))

        dev = torch.device("cuda", self.ctx.device)
        self.w1 = torch.nn.Parameter(kaiming((NCH, 3, 3), 9).to(dev))
        self.b1 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        self.w2 = torch.nn.Parameter(kaiming((NCH, 3, 3, NCH), 9 * NCH).to(dev))
        self.b2 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        self.w3 = torch.nn.Parameter(kaiming((NCH, 3, 3, NCH), 9 * NCH).to(dev))
        self.b3 = torch.nn.Parameter(torch.zeros(NCH, device=dev))
        # the last layer starts at zero: the network begins as the identity
        self.w4 = torch.nn.Parameter(torch.zeros((3, 3, NCH), device=dev))
        self.b4 = torch.nn.Parameter(torch.zeros(1, device=dev))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        squeeze = x.dim() == 2
        xb = x.unsqueeze(0) if squeeze else x
        y = _Correction.apply(xb, self.w1, self.b1, self.w2, self.b2, self.w3, self.b3, self.w4,
                              self.b4, self.ctx)
        return y.squeeze(0) if squeeze else y

    def save(self, path: str, **meta) -> None:
        arrs = {k: v.detach().cpu().numpy() for k, v in self.state_dict().items()}
        arrs.update({f"meta_{k}": np.asarray(v) for k, v in meta.items()})
        np.savez(path, **arrs)

    def load(self, path: str) -> dict:
        z = np.load(path, allow_pickle=False)
        sd = {k: torch.from_numpy(z[k]) for k in z.files if not k.startswith("meta_")}
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
    ax = geom.fp(x.contiguous())
    B = x.shape[0] if x.dim() == 3 else 1
    d = (ax - b).reshape(B, -1).double()
    return d.norm(dim=1) / b.reshape(B, -1).double().norm(dim=1)


# ---------------------------------------------------------------------------- model
class LearnedSIRT:
    """``stages`` x (SIRT step, correction) with the plain-SIRT fallback."""

    def __init__(self, geom: Geometry, net: Correction, stages: int = 5):
        self.geom, self.net, self.stages = geom, net, stages
        self.w = SirtWeights(geom)

    def sirt_step(self, x, b):
        return _SirtStep.apply(x, b, self.geom, self.w.rowinv, self.w.colinv)

    def unrolled(self, b: torch.Tensor, checkpoint: bool = True, all_stages: bool = False):
        """Differentiable forward used for training: images [B, N, N] after the last stage
        (or the list of every stage's output)."""
        from torch.utils.checkpoint import checkpoint as ck
        B = b.shape[0]
        x = torch.zeros((B, self.geom.N, self.geom.N), device=b.device)
        outs = []
        for _ in range(self.stages):
            if checkpoint and torch.is_grad_enabled():
                x = ck(lambda xx, bb: self.net(self.sirt_step(xx, bb)), x, b, use_reentrant=False)
            else:
                x = self.net(self.sirt_step(x, b))
            outs.append(x)
        return outs if all_stages else x

    @torch.no_grad()
    def reconstruct(self, b: torch.Tensor):
        """Returns (images, fell_back [B] bool)."""
        B = b.shape[0]
        x = torch.zeros((B, self.geom.N, self.geom.N), device=b.device)
        xp = None
        for k in range(self.stages):
            xs = self.sirt_step(x, b)
            # both paths start from zero, so plain SIRT's first step is the first stage's (the
            # kernels are deterministic: the same bits)
            xp = xs.clone() if k == 0 else self.sirt_step(xp, b)
            xc = self.net(xs)
            bad = ~torch.isfinite(xc).reshape(B, -1).all(dim=1)
            x = torch.where(bad.view(B, 1, 1), xs, xc)
        worse = residual(self.geom, x, b) > residual(self.geom, xp, b)
        return torch.where(worse.view(B, 1, 1).to(x.device), xp, x), worse
