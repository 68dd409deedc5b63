"""The backprojector's exact per-tile step on inputs that expose single candidates.

A sinogram with one angle gives every pixel at most two candidates, so the per-pixel error is the
per-candidate error itself: these tests hold fast mode to rel-L2 <= 1e-6 and <= 1e-5 per pixel
against the oracle (backprojection.fut:36-93), which the exact step meets (its fraction is the
reference's up to a few roundings) and which a fraction off by the rounding of x_bot + slope
(~1e-4 per candidate at 1024^2) would not.  One angle is chosen so that its |slope| is a rounding
tie in binade [512, 1024): the tiles there take it in their separate tie pass."""
import numpy as np
import pytest
import torch

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# f32 angle whose slope 0.309356689453125 = 5068.5 * 2^-14: halfway between two multiples of
# u = 2^(9-23), the grid of binade 9 (found by scanning f32 angles from 0.3 with libm's sinf/cosf,
# as the host builds its line table)
TIE_ANGLE = np.float32(0.30001866817474365)


def _check(ctx, N, angles):
    R = int(np.ceil(np.sqrt(2) * N))
    rz = -(R - 1) / 2.0
    rng = np.random.default_rng(N + len(angles))
    y = rng.random(len(angles) * R).astype(np.float32)
    ref = oracle.bp(angles, rz, 1.0, N, y)
    ctx.exact = False
    geo = ctx.geometry(angles, rz, 1.0, R, N)
    big = np.abs(ref) > 0.01 * float(np.abs(ref).max())
    for bv in (0, 2):  # 16x16 and 16x64 tiles
        geo.set_variants(-1, bv)
        got = geo.bp(torch.from_numpy(y).cuda()).cpu().numpy().ravel()
        worst = float(np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big])))
        err = rel_l2(got, ref)
        print(f"N={N} angles={list(np.round(angles, 4))} bp variant {bv}: rel-L2 {err:.2e}, worst {worst:.2e}")
        assert err <= 1e-6 and worst <= 1e-5
    ctx.exact = True
    try:
        ex = ctx.geometry(angles, rz, 1.0, R, N).bp(torch.from_numpy(y).cuda()).cpu().numpy().ravel()
    finally:
        ctx.exact = False
    np.testing.assert_array_equal(ex, ref)


@pytest.mark.parametrize("theta", [0.05, 0.4, 0.7853982, 1.1, 1.9, 2.6])
def test_single_angle_backprojection(ctx, theta):
    _check(ctx, 1024, np.array([theta], np.float32))


def test_step_tie_line(ctx):
    _check(ctx, 2048, np.array([TIE_ANGLE], np.float32))
    _check(ctx, 2048, np.array([0.9, TIE_ANGLE, 2.2], np.float32))


@pytest.mark.parametrize("N,theta", [(1500, 0.35), (1500, 2.0), (4096, 0.6), (4096, 1.3)])
def test_single_angle_other_sizes(ctx, N, theta):
    """Image sizes whose tile edges fall off the powers of two (1500: h = 750) and the largest
    sweep size (4096: binades up to [2048, 4096))."""
    _check(ctx, N, np.array([theta], np.float32))

