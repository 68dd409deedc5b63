"""Measured detector data to the sinogram the projector expects (HIP kernels in
csrc/sk_prep.hip).

The reference reconstructs from sinograms that are already line integrals (its SIRT
benchmark reads data/tooth.in, futhark/SIRT.fut:1-2); these are the steps before that:

* ``normalize(raw, flat, dark)``: flat/dark-field normalisation and -log,
  b = -log(max((raw - dark) / (flat - dark), min_transmission)) / scale, with the open-beam
  (flat) and dark frames averaged per detector bin.
* ``rotation_centre(sino, angles)``: the detector position c0 (in bins) of the rotation axis,
  from the centre of mass of every projection, which for a parallel-beam scan traces
  c0 + p cos(theta) + q sin(theta); ``centred_rhozero(c0, deltarho)`` = -c0 * deltarho is the
  rhozero that puts the axis at rho = 0, so a mis-centred scan is reconstructed by passing
  that rhozero to the geometry (no resampling of the data).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from .projector import Context, _ptr, _stream


def _dev(a, ctx: Context) -> torch.Tensor:
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    if t.dtype != torch.float32:
        raise TypeError("expected float32 data")
    return t.to(torch.device("cuda", ctx.device)).contiguous()


def normalize(raw, flat, dark=None, min_transmission: float = 1e-6, scale: float = 1.0,
              out: torch.Tensor | None = None, ctx: Context | None = None) -> torch.Tensor:
    """raw counts [..., A, R] -> line integrals of the same shape (float32, on the GPU).

    flat: open-beam frames [R] or [n_flat, R]; dark: dark frames [R] or [n_dark, R], or None."""
    ctx = ctx or Context.default(0)
    raw = _dev(raw, ctx)
    if raw.dim() < 2:
        raise ValueError("raw must be [..., A, R]")
    A, R = raw.shape[-2], raw.shape[-1]
    batch = raw.numel() // (A * R)
    flat = _dev(flat, ctx).reshape(-1, R) if flat is not None else None
    if flat is None or flat.shape[0] < 1:
        raise ValueError("need at least one flat frame of R bins")
    dark_t = _dev(dark, ctx).reshape(-1, R) if dark is not None else None
    nd = 0 if dark_t is None else dark_t.shape[0]
    if out is None:
        out = torch.empty_like(raw)
    elif out.shape != raw.shape or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor shaped like raw")
    nat.check(nat.lib().sk_normalize(ctx.handle, batch, A, R, _ptr(raw), flat.shape[0], _ptr(flat), nd,
                                     _ptr(dark_t) if nd else None, float(min_transmission),
                                     float(scale), _ptr(out), _stream(ctx.device)),
              ctx.handle, "normalize")
    return out


def row_centroids(sino, ctx: Context | None = None) -> torch.Tensor:
    """Centre of mass (in bins, FP64) of every row of a sinogram [..., A, R] -> [..., A]."""
    ctx = ctx or Context.default(0)
    s = _dev(sino, ctx)
    R = s.shape[-1]
    rows = s.numel() // R
    com = torch.empty(s.shape[:-1], dtype=torch.float64, device=s.device)
    nat.check(nat.lib().sk_row_centroid(ctx.handle, ctypes.c_int64(rows), R, _ptr(s), _ptr(com),
                                        _stream(ctx.device)), ctx.handle, "row centroid")
    return com


def fit_centre(centroids, angles) -> float:
    """Least-squares c0 of centroids ~ c0 + p cos(theta) + q sin(theta) (FP64, host); rows
    with an undefined centroid (empty projection) are left out."""
    c = np.asarray(centroids, np.float64).ravel()
    th = np.asarray(angles, np.float64).ravel()
    if c.size != th.size:
        raise ValueError("one centroid per angle")
    ok = np.isfinite(c)
    if ok.sum() < 3:
        raise ValueError("need at least three projections with mass")
    M = np.stack([np.ones(ok.sum()), np.cos(th[ok]), np.sin(th[ok])], axis=1)
    coef, *_ = np.linalg.lstsq(M, c[ok], rcond=None)
    return float(coef[0])


def rotation_centre(sino, angles, ctx: Context | None = None) -> float:
    """Detector position (bins, fractional) of the rotation axis of one sinogram [A, R]; the
    sinogram should hold non-negative line integrals."""
    com = row_centroids(sino, ctx)
    if com.dim() != 1:
        raise ValueError("one sinogram [A, R] at a time")
    return fit_centre(com.cpu().numpy(), angles)


def centred_rhozero(c0: float, deltarho: float = 1.0) -> float:
    """rhozero that maps detector bin c0 to rho = 0 (rho = rhozero + r * deltarho)."""
    return -float(c0) * float(deltarho)
