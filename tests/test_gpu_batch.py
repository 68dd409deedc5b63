"""Kernel variants, the autotuner and batched persistent SIRT."""
import ctypes
import json
import numpy as np
import pytest
import torch

import oracle
from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu


def _uniform(A):
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


@pytest.mark.parametrize("N,A,R,rz,dr", [(64, 90, 91, -45.0, 1.0), (512, 360, 725, -362.0, 1.0),
                                         (1024, 720, 1449, -724.0, 1.0), (102, 37, 300, -180.5, 0.9),
                                         (40, 23, 30, -29.0, 2.0), (24, 19, 70, -17.25, 0.5)])
def test_all_variants_bitwise_identical(ctx, N, A, R, rz, dr):
    """Every kernel shape, the staged ones included (rays far outside the image on both
    sides in the off-centre case), gives the same bits."""
    geo = ctx.geometry(_uniform(A), rz, dr, R, N)
    x = torch.rand(2, N, N, device="cuda")
    y = torch.rand(2, A, R, device="cuda")
    geo.set_variants(0, 0)
    f0, b0 = geo.fp(x), geo.bp(y)
    for fv in range(8):
        for bv in (0, 1, 2):
            geo.set_variants(fv, bv)
            torch.testing.assert_close(geo.fp(x), f0, rtol=0, atol=0)
            torch.testing.assert_close(geo.bp(y), b0, rtol=0, atol=0)
    geo.set_variants(-1, -1)
    if N == 64:
        ref = oracle.fp(_uniform(A), rz, dr, R, x[0].cpu().numpy())
        assert rel_l2(f0[0].cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("N,A,R,rz,dr", [(512, 360, 725, -362.0, 1.0), (102, 37, 300, -180.5, 0.9)])
def test_sirt_every_fp_variant_bitwise(ctx, N, A, R, rz, dr):
    """SIRT iterations (the residual form of A.x, fused) give the same bits with every forward
    kernel shape; the angles in a shuffled order (the grouped kernel sorts its line groups)."""
    ang = _uniform(A)[np.random.default_rng(5).permutation(A)]
    geo = ctx.geometry(ang, rz, dr, R, N)
    sino = torch.rand(A, R, device="cuda")
    res = {}
    for fv in (0, 6, 7):
        geo.set_variants(fv, 2)
        img = torch.zeros(N, N, device="cuda")
        geo.sirt_(img, sino, 3)
        res[fv] = img
    geo.set_variants(-1, -1)
    for fv in (6, 7):
        torch.testing.assert_close(res[fv], res[0], rtol=0, atol=0)


@pytest.mark.parametrize("N,A,R,rz,dr", [(64, 90, 91, -45.0, 1.0), (40, 23, 30, -29.0, 2.0)])
def test_exact_mode_every_variant_bitwise_oracle(exact, N, A, R, rz, dr):
    """Exact mode is bitwise the oracle with every kernel shape and grid order."""
    rng = np.random.default_rng(8)
    x = rng.random(N * N).astype(np.float32)
    y = rng.random(A * R).astype(np.float32)
    fref, bref = oracle.fp(_uniform(A), rz, dr, R, x), oracle.bp(_uniform(A), rz, dr, N, y)
    geo = exact.geometry(_uniform(A), rz, dr, R, N)
    for v in range(8):
        geo.set_variants(v, min(v, 2))
        np.testing.assert_array_equal(geo.fp(torch.from_numpy(x).cuda()).cpu().numpy().ravel(), fref)
        np.testing.assert_array_equal(geo.bp(torch.from_numpy(y).cuda()).cpu().numpy().ravel(), bref)
    geo.set_variants(-1, -1)


def test_autotune_cache(ctx, tmp_path, monkeypatch):
    from synkrotomo_amd import autotune
    monkeypatch.setenv("SYNKRO_TUNE_CACHE", str(tmp_path / "tune.json"))
    geo = ctx.geometry(_uniform(180), -181.0, 1.0, 363, 256)
    fv, bv = autotune.tuned(geo)
    assert fv in range(8) and bv in (0, 1, 2)
    assert (tmp_path / "tune.json").exists()
    assert autotune.tuned(geo) == (fv, bv)  # served from the cache
    # batches are tuned and cached apart: batch 20 is timed on 32 images
    fb, bb = autotune.tuned(geo, batch=20)
    assert fb in range(8) and bb in (0, 1, 2)
    keys = json.load(open(tmp_path / "tune.json"))
    assert sorted(k.rsplit("/", 1)[1] for k in keys) == ["b1", "b32"]
    assert autotune.tuned(geo, batch=17) == (fb, bb)
    assert [autotune.batch_bucket(b) for b in (1, 2, 3, 16, 17, 64, 500)] == [1, 2, 4, 16, 32, 64, 64]
    geo.set_variants(-1, -1)


def test_autotune_batch_rejects_bad_batch(ctx):
    from synkrotomo_amd import _native as nat
    geo = ctx.geometry(_uniform(12), -11.5, 1.0, 24, 16)
    fv, bv = ctypes.c_int(), ctypes.c_int()
    for bad in (0, -3, 1025):
        rc = nat.lib().sk_autotune_batch(ctx.handle, geo.handle, bad, ctypes.byref(fv), ctypes.byref(bv), None)
        assert rc != 0


@pytest.mark.parametrize("fp_mode", ["", "4", "2"], ids=["grouped", "staged", "per-wave"])
@pytest.mark.parametrize("variants", [(-1, -1), (2, 2)])
@pytest.mark.parametrize("specs", [
    [(32, 30, 46, -22.5, 3), (64, 45, 91, -45.0, 2), (48, 20, 68, -33.5, 4)],
    [(256, 180, 363, -181.0, 2), (512, 360, 725, -362.0, 1)],
], ids=["small", "256-512"])
def test_sirt_batched_equals_per_problem(ctx, specs, variants, fp_mode, monkeypatch):
    """Uneven groups in one persistent launch sequence == each problem run on its own, with
    the per-problem runs on the default and on other kernel shapes than the batched ones, and
    the batched forward items of each kind (grouped staged by default)."""
    from synkrotomo_amd.projector import sirt_batched
    if fp_mode:
        monkeypatch.setenv("SYNKRO_BATCH_FP_MODE", fp_mode)
    gen = torch.Generator(device="cuda").manual_seed(7)
    groups, singles = [], []
    for N, A, R, rz, b in specs:
        geo = ctx.geometry(_uniform(A), rz, 1.0, R, N)
        geo.set_variants(*variants)
        sino = torch.rand(b, A, R, device="cuda", generator=gen)
        img = torch.zeros(b, N, N, device="cuda")
        groups.append((geo, img, sino))
        ref = torch.zeros(b, N, N, device="cuda")
        geo.sirt_(ref, sino, 7)
        singles.append(ref)
    sirt_batched(groups, 7)
    for (geo, img, _), ref in zip(groups, singles):
        torch.testing.assert_close(img, ref, rtol=0, atol=0)


def test_sirt_batched_exact_matches_oracle(exact):
    from synkrotomo_amd.projector import sirt_batched
    g = load_golden("c1.npz")
    geo = exact.geometry(g["angles"], -45.0, 1.0, 91, 64)
    geo2 = exact.geometry(_uniform(12), -11.5, 1.0, 24, 16)
    b2 = np.random.default_rng(4).random((3, 12, 24)).astype(np.float32)
    img1 = torch.zeros(1, 64, 64, device="cuda")
    img2 = torch.zeros(3, 16, 16, device="cuda")
    sirt_batched([(geo, img1, torch.from_numpy(g["fp"]).cuda().reshape(1, 90, 91)),
                  (geo2, img2, torch.from_numpy(b2).cuda())], 10)
    np.testing.assert_array_equal(img1.cpu().numpy().ravel(), g["sirt10"])
    for i in range(3):
        ref = oracle.sirt(_uniform(12), -11.5, 1.0, np.zeros(256, np.float32), b2[i], 10)
        np.testing.assert_array_equal(img2[i].cpu().numpy().ravel(), ref)


def test_sirt_batched_mixed_bin_widths(ctx):
    """Groups with deltarho != 1 take the general-candidate persistent kernel."""
    from synkrotomo_amd.projector import sirt_batched
    geo_a = ctx.geometry(_uniform(19), -17.25, 0.5, 70, 24)
    geo_b = ctx.geometry(_uniform(30), -22.5, 1.0, 46, 32)
    ya = torch.rand(2, 19, 70, device="cuda")
    yb = torch.rand(1, 30, 46, device="cuda")
    xa, xb = torch.zeros(2, 24, 24, device="cuda"), torch.zeros(1, 32, 32, device="cuda")
    sirt_batched([(geo_a, xa, ya), (geo_b, xb, yb)], 5)
    for i in range(2):
        ref = oracle.sirt(_uniform(19), -17.25, 0.5, np.zeros(576, np.float32), ya[i].cpu().numpy(), 5)
        assert rel_l2(xa[i].cpu().numpy().ravel(), ref) <= 1e-5
    ref = oracle.sirt(_uniform(30), -22.5, 1.0, np.zeros(1024, np.float32), yb[0].cpu().numpy(), 5)
    assert rel_l2(xb[0].cpu().numpy().ravel(), ref) <= 1e-5
