"""The reference's comparison tools (futhark/tools/*.fut) on the GPU: ``sk_compare``."""
from __future__ import annotations

import ctypes

import torch

from . import _native as nat
from .projector import Context, _ptr, _stream

REFERENCE_TOL = 0.05  # tools/sanityCheck.fut:2, tools/diff.fut:2


def compare(sane: torch.Tensor, test: torch.Tensor, tol: float = REFERENCE_TOL,
            ctx: Context | None = None) -> tuple[bool, int, float]:
    """(all |sane - test| < tol, count of indices i > 0 with |sane - test| > tol,
    mean of the f32 squared differences) for two float32 device tensors of equal size."""
    if sane.dtype != torch.float32 or test.dtype != torch.float32:
        raise TypeError("compare needs float32 tensors")
    if sane.numel() != test.numel():
        raise ValueError(f"sizes differ: {sane.numel()} vs {test.numel()}")
    if sane.numel() == 0:
        raise ValueError("empty arrays")
    ctx = ctx or Context.default(sane.device.index or 0)
    a, b = sane.contiguous(), test.contiguous()
    ok, over, mse = ctypes.c_int(), ctypes.c_int64(), ctypes.c_float()
    nat.check(nat.lib().sk_compare(ctx.handle, a.numel(), _ptr(a), _ptr(b), float(tol),
                                   ctypes.byref(ok), ctypes.byref(over), ctypes.byref(mse),
                                   _stream(ctx.device)), ctx.handle, "compare")
    return bool(ok.value), int(over.value), float(mse.value)
