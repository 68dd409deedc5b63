"""Parity at the full sizes the benchmarks run.

* 2048^2 image, 2048 angles in [0, pi), 2897 bins, rhozero -1448 (the GUPS configuration):
  the whole sinogram of A.x and the whole image of B.r against the oracle
  (forwardprojection.fut:71-84, backprojection.fut:82-93), with every kernel variant -- bitwise
  in exact mode, rel-L2 <= 1e-5 in fast mode;
* the explicit-matrix SpMV at the same geometry (forwardprojection_sparse.fut:26-40,
  backprojection_sparse.fut:26-37), on the SELL-64 kernel with 16-bit column steps and the
  row-slot kernel: nnz = 1.09e10 > 2^31, so the 64-bit slice offsets are exercised;
* the batched-SIRT problem set (64 x 256^2/180, 96 x 512^2/360, 32 x 1024^2/720 and a SIRT3D
  stack of 512 slices of 512^2/512) through one batched launch sequence: bitwise equal to the
  groups run one by one, and one problem of each size against the oracle's SIRT
  (SIRT.fut:9-49, SIRT3D.fut:18-28).

The oracle runs on the host's cores (about 16 s for the 2048^2 pair on 16 cores).
"""
import numpy as np
import pytest
import torch

import oracle
from conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5

N, A, R, RZ = 2048, 2048, 2897, -1448.0


def _uniform(a):
    return (np.arange(a, dtype=np.float64) * np.pi / a).astype(np.float32)


@pytest.fixture(scope="module")
def c2():
    """Seeded uniform image and sinogram and the oracle's fp / bp of them (all host cores)."""
    rng = np.random.default_rng(2048)
    x = rng.random(N * N, dtype=np.float32)
    y = rng.random(A * R, dtype=np.float32)
    ang = _uniform(A)
    nt = oracle_threads()
    print(f"oracle at 2048^2 on {nt} threads", flush=True)
    return {"ang": ang, "x": x, "y": y,
            "fx": oracle.fp(ang, RZ, 1.0, R, x, nthreads=nt),
            "by": oracle.bp(ang, RZ, 1.0, N, y, nthreads=nt)}


def _run_dense(ctx, c2, fv, bv):
    geo = ctx.geometry(c2["ang"], RZ, 1.0, R, N)
    geo.set_variants(fv, bv)
    try:
        xd = torch.from_numpy(c2["x"]).cuda()
        yd = torch.from_numpy(c2["y"]).cuda()
        fx = geo.fp(xd).cpu().numpy().ravel()
        by = geo.bp(yd).cpu().numpy().ravel()
    finally:
        geo.set_variants(-1, -1)
    return fx, by


@pytest.mark.parametrize("fv,bv", [(0, 0), (1, 1), (2, 2), (3, 0), (4, 1), (5, 2), (6, 2)])
def test_c2_dense_whole_image_exact_bitwise(exact, c2, fv, bv):
    """Exact mode: every ray of the 5.9 M-entry sinogram and every pixel of the 4.2 M-pixel
    image bitwise equal to the oracle, for each kernel shape and grid order (the staged
    variants 5 and 6 run the per-wave kernel in exact mode)."""
    fx, by = _run_dense(exact, c2, fv, bv)
    np.testing.assert_array_equal(fx, c2["fx"])
    np.testing.assert_array_equal(by, c2["by"])


@pytest.mark.parametrize("fv,bv", [(0, 0), (1, 1), (2, 2), (3, 0), (4, 1), (5, 2), (6, 2)])
def test_c2_dense_whole_image_fast(ctx, c2, fv, bv):
    """Fast mode (hardware reciprocal in the step quotients), every kernel variant including
    the tuned pair at this geometry (staged fp 6, 16x64 bp tiles): rel-L2 over the whole
    sinogram and image, and the largest per-entry relative error on entries above 1% of the
    maximum (a tile- or block-boundary fault would show there even when rel-L2 hides it)."""
    fx, by = _run_dense(ctx, c2, fv, bv)
    out = []
    for got, ref in ((fx, c2["fx"]), (by, c2["by"])):
        err = rel_l2(got, ref)
        big = np.abs(ref) > 0.01 * float(np.abs(ref).max())
        worst = float(np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big])))
        zero_ok = bool(np.all(got[ref == 0] == 0))
        out.append((err, worst, zero_ok))
    print(f"variants fp {fv} bp {bv}: fp rel-L2 {out[0][0]:.2e} max {out[0][1]:.2e}; "
          f"bp rel-L2 {out[1][0]:.2e} max {out[1][1]:.2e}")
    for err, worst, zero_ok in out:
        assert err <= TOL
        assert worst <= 1e-4
        assert zero_ok


@pytest.mark.parametrize("kind,variant", [("fp", 9), ("fp", 10), ("bp", 9), ("bp", 10)])
def test_c2_sparse_spmv_tuned_variant(ctx, c2, kind, variant):
    """The explicit matrix at 2048^2 (A: 5.9 M rows; B: 4.2 M rows; 1.09e10 entries each, so
    slice offsets and row pointers run past 2^31) on the SpMV kernels the autotuner picks
    there: variant 9 (SELL-64, 16-bit column steps) and variant 10 (row slots): the whole output
    against the oracle and against the dense kernel."""
    geo = ctx.geometry(c2["ang"], RZ, 1.0, R, N)
    inp = torch.from_numpy(c2["x"] if kind == "fp" else c2["y"]).cuda()
    dense = (geo.fp(inp) if kind == "fp" else geo.bp(inp)).cpu().numpy().ravel()
    m = geo.sparse(kind)
    try:
        assert m.nnz > 2 ** 31
        m.set_spmv_variant(variant)
        got = m.spmv(inp).cpu().numpy().ravel()
        assert m.spmv_variant == variant
    finally:
        geo.drop_sparse()
        torch.cuda.empty_cache()
    ref = c2["fx"] if kind == "fp" else c2["by"]
    e_ref, e_dense = rel_l2(got, ref), rel_l2(got, dense)
    print(f"sparse {kind} variant {variant} nnz {m.nnz}: vs oracle {e_ref:.2e}, vs dense {e_dense:.2e}")
    assert e_ref <= TOL and e_dense <= TOL


C3 = [(256, 180, 64), (512, 360, 96), (1024, 720, 32), (512, 512, 512)]


def test_c3_batched_problem_set(ctx):
    """The batched problem set, 3 iterations: batched == per-group SIRT bitwise for all 704
    images (each geometry on its autotuned kernels), and problem 0 of each group within 1e-5
    of the oracle's SIRT."""
    from synkrotomo_amd.autotune import tuned
    from synkrotomo_amd.projector import centred_bins, sirt_batched
    iters = 3
    gen = torch.Generator(device="cuda").manual_seed(30)
    groups, singles = [], []
    for n, a, cnt in C3:
        r, rz = centred_bins(n)
        geo = ctx.geometry(_uniform(a), rz, 1.0, r, n)
        tuned(geo)
        sino = torch.rand(cnt, a, r, device="cuda", generator=gen)
        groups.append((geo, torch.zeros(cnt, n, n, device="cuda"), sino))
    print("batched SIRT: 64x256^2/180 + 96x512^2/360 + 32x1024^2/720 + 512x512^2/512", flush=True)
    sirt_batched(groups, iters)
    for geo, img, sino in groups:
        ref = torch.zeros_like(img)
        geo.sirt_(ref, sino, iters)
        singles.append(ref)
    torch.cuda.synchronize()
    for (geo, img, _), ref in zip(groups, singles):
        assert torch.equal(img, ref), f"batched != per-group at N={geo.N}, A={geo.A}"
    nt = oracle_threads()
    for (geo, img, sino), (n, a, cnt) in zip(groups, C3):
        # SIRT3D's slice k is SIRT on slab k from a zero image (the oracle's sirt3d is pinned
        # to that in tests/test_oracle.py), so the stack's first slice checks against sirt.
        got = img[0].cpu().numpy().ravel()
        b0 = sino[0].cpu().numpy().ravel()
        ref = oracle.sirt(geo.angles, geo.rho0, geo.drho, np.zeros(n * n, np.float32), b0, iters, nthreads=nt)
        err = rel_l2(got, ref)
        print(f"N={n} A={a}: batched SIRT vs oracle rel-L2 {err:.2e}")
        assert err <= TOL
    del groups, singles
    torch.cuda.empty_cache()
