"""The backprojector's exact-step path (tile_binade / binade_step in csrc/sk_kern.h), restated
in numpy float32 and checked against the reference's own expressions (backprojection.fut:55-62):
inside a tile whose candidate range lies in one binade, every candidate that can have a nonzero
length has |fl(fl(x_bot + slope) - x_bot)| equal to the tile's step, and the moved pixel edges
give the reference's x_top - xmin (slope >= 0) and xmin + 1 - x_top (slope < 0) bit for bit."""
import math

import numpy as np
import pytest

import oracle

f32 = np.float32


def tile_binade(x0, x1):
    lo, hi = x0 - 4, x1 + 5
    if lo <= 0 <= hi:
        return -1
    a, b = (lo, hi) if lo > 0 else (-hi, -lo)
    ea, eb = a.bit_length() - 1, b.bit_length() - 1
    return ea if ea == eb else -1


def binade_step(slope, e):
    t = f32(math.ldexp(abs(float(slope)), 23 - e))
    ft = np.floor(t)
    if t - ft == f32(0.5):
        return f32(0)
    return f32(math.ldexp(float(np.rint(t)), e - 23))


@pytest.mark.parametrize("N,A,tw,th", [(512, 360, 16, 64), (2048, 128, 16, 64), (256, 90, 16, 16),
                                        (2048, 128, 16, 16)])
def test_exact_step_matches_reference_rounding(N, A, tw, th):
    R = int(math.ceil(math.sqrt(2) * N))
    rho0 = f32(-(R - 1) / 2)
    ang = (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)
    c, s, _, _, ns = oracle.preprocess(ang)
    h = N // 2
    fact = f32(np.sqrt(f32(2)) / f32(2))
    checked = tiles = 0
    for j in range(A):
        cost, sint = f32(c[j]), f32(s[j])
        slope = f32(-sint / cost)
        if abs(slope) < 1e-3:
            continue
        span = th if j >= ns else tw
        for t0 in range(0, N, span):
            e = tile_binade(t0 - h, t0 - h + span - 1)
            if e < 0:
                continue
            d = binade_step(slope, e)
            if d == 0:
                continue
            tiles += 1
            xm = np.arange(t0 - h, t0 - h + span, dtype=np.float32)[None, :]
            ym = np.arange(-h, h, 4, dtype=np.float32)[:, None]
            XM, YM = (v.astype(f32) for v in np.broadcast_arrays(xm, ym))
            px = f32(XM + f32(0.5)) - f32(fact * cost)
            py = f32(YM + f32(0.5)) - f32(fact * sint)
            rho = f32(f32(cost * px) + f32(sint * py))
            sf = np.ceil(f32(rho - rho0)).astype(f32)
            xbase = f32(YM * slope)
            x1 = f32(XM + f32(1))
            for k in range(2):
                inter = f32(f32(f32(sf + f32(k)) + rho0) / cost)
                xb = f32(xbase + inter)
                xt = f32(xb + slope)
                near = np.abs(xb - XM) < 3
                checked += int(near.sum())
                np.testing.assert_array_equal(np.abs(f32(xt - xb))[near], np.full(int(near.sum()), d))
                if slope >= 0:
                    np.testing.assert_array_equal(f32(xt - XM)[near], f32(xb - f32(XM - d))[near])
                else:
                    np.testing.assert_array_equal(f32(x1 - xt)[near], f32(f32(x1 + d) - xb)[near])
    assert tiles > 0 and checked > 10000


def test_step_ties_are_refused():
    """|slope| exactly halfway between two multiples of u = 2^(e-23): round-half-even makes
    |fl(x_bot + slope) - x_bot| depend on x_bot's last bit, so the tile must not use one step."""
    e = 9
    u = math.ldexp(1.0, e - 23)
    slope = f32(0.3 - (0.3 % u) + u / 2)  # an exact tie
    assert binade_step(slope, e) == 0
    xb0 = f32(600.0)  # in binade 9: [512, 1024)
    xb1 = np.nextafter(xb0, f32(1e9))
    d0 = f32(f32(xb0 + slope) - xb0)
    d1 = f32(f32(xb1 + slope) - xb1)
    assert d0 != d1
    assert binade_step(np.nextafter(slope, f32(1)), e) > 0
