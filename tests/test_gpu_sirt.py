"""HIP SIRT (fused kernels, hipGraph replay) against the oracle's SIRT.fut restatement."""
import numpy as np
import pytest
import torch

import oracle
from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_sirt_c1_exact_bitwise(exact):
    g = load_golden("c1.npz")
    geo = exact.geometry(g["angles"], -45.0, 1.0, 91, 64)
    x = torch.zeros(64 * 64, device="cuda")
    geo.sirt_(x, _dev(g["fp"]), 10)
    np.testing.assert_array_equal(x.cpu().numpy(), g["sirt10"])


def test_sirt_c1_fast_and_fp64_error(ctx):
    g = load_golden("c1.npz")
    geo = ctx.geometry(g["angles"], -45.0, 1.0, 91, 64)
    x = torch.zeros(64 * 64, device="cuda")
    geo.sirt_(x, _dev(g["fp"]), 10)
    got = x.cpu().numpy()
    err = rel_l2(got, g["sirt10"])
    f64 = oracle.sirt_f64(g["angles"], -45.0, 1.0, np.zeros(64 * 64), g["fp"].astype(np.float64), 10)
    print(f"SIRT c1 fast rel-L2 vs FP32 oracle {err:.3e}; vs FP64 oracle {rel_l2(got, f64):.3e} "
          f"(FP32 oracle vs FP64 {rel_l2(g['sirt10'], f64):.3e})")
    assert err <= TOL


def test_sirt_split_calls_equal_one_call(exact):
    """10 iterations = 3 + 7 (graph block + remainder paths, cached graph reuse)."""
    g = load_golden("c1.npz")
    geo = exact.geometry(g["angles"], -45.0, 1.0, 91, 64)
    b = _dev(g["fp"])
    x1 = torch.zeros(64 * 64, device="cuda")
    geo.sirt_(x1, b, 10)
    x2 = torch.zeros(64 * 64, device="cuda")
    geo.sirt_(x2, b, 3)
    geo.sirt_(x2, b, 7)
    torch.testing.assert_close(x1, x2, rtol=0, atol=0)
    x3 = torch.zeros(64 * 64, device="cuda")
    for _ in range(10):
        geo.sirt_(x3, b, 1)
    torch.testing.assert_close(x1, x3, rtol=0, atol=0)


def test_sirt_weights_match_oracle(exact):
    g = load_golden("c1.npz")
    geo = exact.geometry(g["angles"], -45.0, 1.0, 91, 64)
    rw, cw = geo.sirt_weights()
    with np.errstate(divide="ignore"):
        rs = oracle.fp(g["angles"], -45.0, 1.0, 91, np.ones(64 * 64, np.float32))
        cs = oracle.bp(g["angles"], -45.0, 1.0, 64, np.ones(90 * 91, np.float32))
    np.testing.assert_array_equal(rw.cpu().numpy().ravel(), np.where(rs == 0, 0, 1 / rs).astype(np.float32))
    np.testing.assert_array_equal(cw.cpu().numpy().ravel(), np.where(cs == 0, 0, 1 / cs).astype(np.float32))


def test_sirt_entry_point_and_resume(ctx):
    """SIRT.SIRT().main keeps the caller's image; passing the result back resumes."""
    from synkrotomo_amd.futhark import SIRT
    g = load_golden("c1.npz")
    x0 = np.zeros(64 * 64, np.float32)
    s = SIRT.SIRT()
    r10 = s.main(g["angles"], -45.0, 1.0, x0, g["fp"], 10).get()
    assert np.all(x0 == 0)
    r4 = s.main(g["angles"], -45.0, 1.0, x0, g["fp"], 4).get()
    r10b = s.main(g["angles"], -45.0, 1.0, r4, g["fp"], 6).get()
    assert rel_l2(r10, g["sirt10"]) <= TOL
    assert rel_l2(r10b, r10) <= 1e-6


def test_sirt3d_matches_oracle(exact):
    from synkrotomo_amd.futhark import SIRT3D
    N, A, R = 16, 12, 24
    ang = (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)
    rhos = (np.arange(R, dtype=np.float32) - np.float32(11.5)).astype(np.float32)
    proj = np.random.default_rng(9).random(N * A * R).astype(np.float32)
    out = SIRT3D.SIRT3D().main(ang, rhos, proj, 5, N).get()
    np.testing.assert_array_equal(out, oracle.sirt3d(ang, rhos, proj, 5, N))


def test_sirt_batch_equals_single(ctx):
    N, A, R = 32, 24, 46
    ang = (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)
    geo = ctx.geometry(ang, -22.5, 1.0, R, N)
    b = torch.rand(4, A, R, device="cuda")
    xb = torch.zeros(4, N, N, device="cuda")
    geo.sirt_(xb, b, 6)
    for i in range(4):
        x = torch.zeros(N, N, device="cuda")
        geo.sirt_(x, b[i].contiguous(), 6)
        torch.testing.assert_close(xb[i], x, rtol=0, atol=0)


def test_sirt_512_variant_vs_oracles(ctx):
    """The 512^2 variant of the 2048^2 SIRT configuration: 50-ellipse phantom, 512 angles in
    [0, pi), 725 bins, rhozero -362, sinogram from the FP32 oracle's forward projector, 20
    iterations from a zero image (the oracle needs about 1 s per iteration on 16 cores; the
    2048^2 run itself is timed by bench.py).  GPU (fast mode) vs the FP32 oracle within 1e-5;
    the FP64 oracle's distance to both is printed."""
    from synkrotomo_amd.phantom import ellipse_phantom
    from synkrotomo_amd.projector import uniform_angles
    N, A, R, rz, iters = 512, 512, 725, -362.0, 20
    ang = uniform_angles(A)
    ph = ellipse_phantom(N, 50, seed=512).ravel()
    nt = min(16, oracle.default_threads())
    b = oracle.fp(ang, rz, 1.0, R, ph, nthreads=nt)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    np.testing.assert_allclose(geo.fp(_dev(ph)).cpu().numpy().ravel(), b, rtol=0,
                               atol=1e-5 * float(np.abs(b).max()))
    x = torch.zeros(N * N, device="cuda")
    geo.sirt_(x, _dev(b), iters)
    got = x.cpu().numpy()
    ref32 = oracle.sirt(ang, rz, 1.0, np.zeros(N * N, np.float32), b, iters, nthreads=nt)
    ref64 = oracle.sirt_f64(ang, rz, 1.0, np.zeros(N * N), b.astype(np.float64), iters, nthreads=nt)
    err = rel_l2(got, ref32)
    print(f"SIRT 512^2/512/725, {iters} it: fast vs FP32 oracle {err:.3e}; vs FP64 oracle "
          f"{rel_l2(got, ref64):.3e} (FP32 oracle vs FP64 {rel_l2(ref32, ref64):.3e}); "
          f"image error vs phantom {rel_l2(got, ph):.3e}")
    assert err <= TOL
