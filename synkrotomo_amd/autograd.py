"""Differentiable projector and SIRT steps (torch.autograd.Function on the HIP kernels).

Every operator's backward is the exact adjoint of its forward, each a hand-written kernel:
  project      y = A x        backward  A^T g  (sk_bp_adjoint)
  backproject  x = B r        backward  B^T g  (sk_bp_transpose)
where A is the forward projector (forwardprojection.fut:21-59) and B the reference's
backprojector (backprojection.fut:36-78).  B equals A^T except on lines whose slope rounds
to zero (theta = 0, pi/2), where the reference drops rays lying on pixel edges; keeping the
two adjoints separate makes every gradient exact.

sirt_step   x' = x + C * B(R * (b - A x))   (SIRT.fut:27-39), affine in (x, b):
            dx = g - A^T(R * B^T(C * g)),   db = R * B^T(C * g).
The step saves nothing for backward (its Jacobian does not depend on x), so an unrolled
chain of steps costs O(1) activation memory; ``unrolled_sirt`` additionally accepts a
per-step module (e.g. a learned correction) and checkpoints each segment with
torch.utils.checkpoint so only segment boundaries are kept.
"""
from __future__ import annotations

import torch
from torch.utils.checkpoint import checkpoint

from .projector import Geometry


def _c(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_contiguous() else t.contiguous()


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, geom: Geometry):
        ctx.geom = geom
        return geom.fp(_c(x))

    @staticmethod
    def backward(ctx, g):
        return ctx.geom.bp_adjoint(_c(g)), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, geom: Geometry):
        ctx.geom = geom
        return geom.bp(_c(r))

    @staticmethod
    def backward(ctx, g):
        return ctx.geom.bp_transpose(_c(g)), None


class _SirtStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, b, geom: Geometry, rowinv, colinv):
        ctx.geom, ctx.rowinv, ctx.colinv = geom, rowinv, colinv
        ctx.x_shape, ctx.b_shape = x.shape, b.shape
        out = _c(x).clone()
        geom.sirt_step_(out, _c(b), rowinv, colinv)
        return out

    @staticmethod
    def backward(ctx, g):
        geom = ctx.geom
        g2 = _c(g).reshape(-1, geom.N, geom.N)
        t = geom.bp_transpose(_c(g2 * ctx.colinv))      # B^T (C g)   [batch, A, R]
        gb = t * ctx.rowinv                               # R B^T C g
        gx = g2 - geom.bp_adjoint(_c(gb))                 # g - A^T R B^T C g
        return gx.reshape(ctx.x_shape), gb.reshape(ctx.b_shape), None, None, None


def project(x: torch.Tensor, geom: Geometry) -> torch.Tensor:
    """A x for images [..., N, N]; differentiable (backward = exact A^T)."""
    return _Project.apply(x, geom)


def backproject(r: torch.Tensor, geom: Geometry) -> torch.Tensor:
    """B r (the reference backprojection) for sinograms [..., A, R]; backward = exact B^T."""
    return _Backproject.apply(r, geom)


class SirtWeights:
    """R = 1/rowsum(A) [A, R] and C = 1/colsum(B) [N, N] of a geometry, computed once."""

    def __init__(self, geom: Geometry):
        self.rowinv, self.colinv = geom.sirt_weights()


def sirt_step(x: torch.Tensor, b: torch.Tensor, geom: Geometry, weights: SirtWeights | None = None):
    """One differentiable SIRT iteration x + C*B(R*(b - A x)) on [..., N, N] / [..., A, R]."""
    w = weights or SirtWeights(geom)
    return _SirtStep.apply(x, b, geom, w.rowinv, w.colinv)


def unrolled_sirt(x0: torch.Tensor, b: torch.Tensor, geom: Geometry, steps: int,
                  correction=None, checkpoint_segments: bool = True) -> torch.Tensor:
    """``steps`` differentiable SIRT iterations from x0.  If ``correction`` (a callable on
    [batch, N, N] images, e.g. an nn.Module) is given, it is applied after every step; each
    step + correction segment is checkpointed, so backward recomputes it instead of keeping
    its activations."""
    w = SirtWeights(geom)

    def segment(x, bb):
        y = _SirtStep.apply(x, bb, geom, w.rowinv, w.colinv)
        return correction(y) if correction is not None else y

    x = x0
    for _ in range(int(steps)):
        if checkpoint_segments and correction is not None and torch.is_grad_enabled():
            x = checkpoint(segment, x, b, use_reentrant=False)
        else:
            x = segment(x, b)
    return x
