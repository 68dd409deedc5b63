"""Seeded random-ellipse phantoms (the synthetic inputs of the benchmark configurations).

Each phantom is a sum of ``n_ellipses`` filled ellipses on normalised coordinates
[-1, 1]^2 evaluated at pixel centres, clipped to [0, 1]:
  centre U[-0.7, 0.7]^2, semi-axes U[0.05, 0.4]^2, rotation U[0, pi), intensity U[-0.5, 1].
Draws come from ``numpy.random.default_rng(seed)`` in that order, one array per parameter.
Row i of the image is y = (2i+1)/N - 1, column j is x = (2j+1)/N - 1.
"""
from __future__ import annotations

import numpy as np


def ellipse_params(rng: np.random.Generator, n_ellipses: int):
    centres = rng.uniform(-0.7, 0.7, size=(n_ellipses, 2))
    axes = rng.uniform(0.05, 0.4, size=(n_ellipses, 2))
    rot = rng.uniform(0.0, np.pi, size=n_ellipses)
    val = rng.uniform(-0.5, 1.0, size=n_ellipses)
    return centres, axes, rot, val


def render(N: int, centres, axes, rot, val) -> np.ndarray:
    c = (2.0 * np.arange(N) + 1.0) / N - 1.0
    x = c[None, :]
    y = c[:, None]
    img = np.zeros((N, N), np.float64)
    for (cx, cy), (a, b), t, v in zip(centres, axes, rot, val):
        dx, dy = x - cx, y - cy
        u = (dx * np.cos(t) + dy * np.sin(t)) / a
        w = (-dx * np.sin(t) + dy * np.cos(t)) / b
        img += np.where(u * u + w * w <= 1.0, v, 0.0)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def ellipse_phantom(N: int, n_ellipses: int = 10, seed: int = 0) -> np.ndarray:
    """One N x N float32 phantom."""
    rng = np.random.default_rng(seed)
    return render(N, *ellipse_params(rng, n_ellipses))


def ellipse_phantoms(count: int, N: int, n_ellipses: int = 10, seed: int = 0) -> np.ndarray:
    """``count`` phantoms [count, N, N] from one generator stream."""
    rng = np.random.default_rng(seed)
    return np.stack([render(N, *ellipse_params(rng, n_ellipses)) for _ in range(count)])
