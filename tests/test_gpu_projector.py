"""HIP projector parity against the oracle (the FP32 restatement of the reference).

Exact mode (SK_OPT_EXACT: IEEE division everywhere) must be bitwise equal to the oracle;
the default fast mode (hardware reciprocal in the per-step quotients) must be within
relative L2 1e-5 (the parity tolerance of the path), and is reported tighter than that.
"""
import hashlib

import numpy as np
import pytest
import torch

import oracle
from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _uniform(A):
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


def _c1(ctx):
    g = load_golden("c1.npz")
    geo = ctx.geometry(g["angles"], -45.0, 1.0, 91, 64)
    return g, geo


def test_fp_c1_exact_bitwise(exact):
    g, geo = _c1(exact)
    out = geo.fp(_dev(g["phantom"])).cpu().numpy().ravel()
    np.testing.assert_array_equal(out, g["fp"])


def test_fp_c1_fast(ctx):
    g, geo = _c1(ctx)
    out = geo.fp(_dev(g["phantom"])).cpu().numpy().ravel()
    err = rel_l2(out, g["fp"])
    print(f"fp c1 fast rel-L2 {err:.3e}")
    assert err <= TOL


def test_bp_c1_exact_bitwise(exact):
    g, geo = _c1(exact)
    out = geo.bp(_dev(g["fp"])).cpu().numpy().ravel()
    np.testing.assert_array_equal(out, g["bp"])


def test_bp_c1_fast(ctx):
    g, geo = _c1(ctx)
    out = geo.bp(_dev(g["fp"])).cpu().numpy().ravel()
    err = rel_l2(out, g["bp"])
    print(f"bp c1 fast rel-L2 {err:.3e}")
    assert err <= TOL


def _fixture512():
    g = load_golden("fpinput512.npz")
    image = (g["image_k"].astype(np.float64) / 255.0).astype(np.float32).ravel()
    return g, image


def test_reference_fixture_512_exact(exact):
    """The reference's own benchmark input (futhark/data/fpinputf32rad512): all 805
    angles flat, so only the transposed branch runs."""
    g, image = _fixture512()
    geo = exact.geometry(g["angles"], g["rhozero"], g["deltarho"], int(g["numrhos"]), 512)
    fp = geo.fp(_dev(image)).cpu().numpy().ravel()
    assert hashlib.sha256(fp.tobytes()).hexdigest() == str(g["fp_sha"])
    bp = geo.bp(_dev(fp)).cpu().numpy().ravel()
    assert hashlib.sha256(bp.tobytes()).hexdigest() == str(g["bp_sha"])


def test_reference_fixture_512_fast(ctx):
    g, image = _fixture512()
    geo = ctx.geometry(g["angles"], g["rhozero"], g["deltarho"], int(g["numrhos"]), 512)
    fp = geo.fp(_dev(image)).cpu().numpy().reshape(-1, int(g["numrhos"]))
    e1 = rel_l2(fp[::8], g["fp_rows"])
    bp = geo.bp(_dev(fp)).cpu().numpy().reshape(512, 512)
    e2 = rel_l2(bp[::8], g["bp_rows"])
    print(f"fixture512 fast rel-L2 fp {e1:.3e} bp {e2:.3e}")
    assert e1 <= TOL and e2 <= TOL


# Edge cases: tiny images, one angle, sparse and unordered angles, angles outside [0, pi),
# off-centre detectors (rays that miss the image), other bin widths (1 and 3 candidates).
EDGE = [
    ("N2", 2, np.array([0.3, 1.9], np.float32), 3, -1.0, 1.0),
    ("A1", 16, np.array([0.7], np.float32), 23, -11.0, 1.0),
    ("sparse5", 48, np.array([2.5, 0.1, 1.2, 0.9, 3.0], np.float32), 68, -33.5, 1.0),
    ("wide_angles", 32, np.array([-2.0, 4.0, 7.5, -0.4, 3.3, 1.5707964], np.float32), 46, -22.5, 1.0),
    ("offcentre", 32, _uniform(17), 40, 5.0, 1.0),
    ("drho_half", 24, _uniform(19), 70, -17.25, 0.5),
    ("drho_two", 40, _uniform(23), 30, -29.0, 2.0),
    ("axis_edges", 64, np.array([0.0, 1.5707964, 0.7853982, 2.3561945], np.float32), 91, -45.0, 1.0),
    # every ray misses the image: all-zero outputs through the far end of the index ranges
    ("far_miss", 32, _uniform(9), 40, 1e5, 1.0),
    # a long detector: bin indices near 5e4 and a 100,003-bin sinogram row
    ("long_detector", 32, _uniform(7), 100003, -50001.0, 1.0),
]


@pytest.mark.parametrize("name,N,ang,R,rz,dr", EDGE, ids=[e[0] for e in EDGE])
def test_edge_cases_exact_bitwise(exact, name, N, ang, R, rz, dr):
    rng = np.random.default_rng(11)
    x = rng.random(N * N).astype(np.float32)
    y = rng.random(ang.size * R).astype(np.float32)
    geo = exact.geometry(ang, rz, dr, R, N)
    np.testing.assert_array_equal(geo.fp(_dev(x)).cpu().numpy().ravel(), oracle.fp(ang, rz, dr, R, x))
    np.testing.assert_array_equal(geo.bp(_dev(y)).cpu().numpy().ravel(), oracle.bp(ang, rz, dr, N, y))


@pytest.mark.parametrize("name,N,ang,R,rz,dr", EDGE, ids=[e[0] for e in EDGE])
def test_edge_cases_fast(ctx, name, N, ang, R, rz, dr):
    rng = np.random.default_rng(11)
    x = rng.random(N * N).astype(np.float32)
    y = rng.random(ang.size * R).astype(np.float32)
    geo = ctx.geometry(ang, rz, dr, R, N)
    assert rel_l2(geo.fp(_dev(x)).cpu().numpy().ravel(), oracle.fp(ang, rz, dr, R, x)) <= TOL
    assert rel_l2(geo.bp(_dev(y)).cpu().numpy().ravel(), oracle.bp(ang, rz, dr, N, y)) <= TOL


def test_batch_equals_single(ctx):
    N, A, R = 32, 30, 46
    ang = _uniform(A)
    geo = ctx.geometry(ang, -22.5, 1.0, R, N)
    xs = torch.rand(3, N, N, device="cuda")
    out = geo.fp(xs)
    for i in range(3):
        torch.testing.assert_close(out[i], geo.fp(xs[i].contiguous()), rtol=0, atol=0)
    ys = torch.rand(3, A, R, device="cuda")
    back = geo.bp(ys)
    for i in range(3):
        torch.testing.assert_close(back[i], geo.bp(ys[i].contiguous()), rtol=0, atol=0)


def test_adjoint_matches_oracle_and_dot_test(ctx):
    N, A, R = 64, 90, 91
    ang = _uniform(A)  # includes theta = 0 and pi/2
    geo = ctx.geometry(ang, -45.0, 1.0, R, N)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(N * N).astype(np.float32)
    y = rng.standard_normal(A * R).astype(np.float32)
    aty = geo.bp_adjoint(_dev(y)).cpu().numpy().ravel()
    assert rel_l2(aty, oracle.fp_adjoint(ang, -45.0, 1.0, R, N, y)) <= TOL
    ax = geo.fp(_dev(x)).cpu().numpy().ravel()
    lhs = float(np.dot(ax.astype(np.float64), y))
    rhs = float(np.dot(x.astype(np.float64), aty.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))


def _sample_rays(ang, R, rng, k):
    lines = oracle.preprocess(ang)
    c, s, l, o, ns = lines
    pick = rng.choice(ang.size, size=min(k, ang.size), replace=False)
    return [(j, int(rng.integers(0, R))) for j in pick], lines


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_full_size_2048_sampled_parity_and_linearity(ctx, mode):
    """2048^2, 2048 angles, 2897 bins: sampled rays and pixels against the oracle, plus
    linearity and the closed-form fp(1)."""
    ctx.exact = mode == "exact"
    try:
        N, A, R, rz = 2048, 2048, 2897, -1448.0
        ang = _uniform(A)
        geo = ctx.geometry(ang, rz, 1.0, R, N)
        rng = np.random.default_rng(4)
        x = torch.rand(N, N, device="cuda")
        fx = geo.fp(x).cpu().numpy()
        xn = x.cpu().numpy()
        xT = np.ascontiguousarray(xn.T)
        (c, s, l, o, ns) = oracle.preprocess(ang)
        worst = 0.0
        for _ in range(48):
            j = int(rng.integers(0, A))
            r = int(rng.integers(0, R))
            frame = xT if j >= ns else xn
            ref = oracle.fp_ray_line(c[j], s[j], l[j], r, rz, 1.0, N // 2, frame)
            got = fx[o[j], r]
            if mode == "exact":
                assert got == ref, (j, r)
            worst = max(worst, abs(got - ref) / max(abs(ref), 1e-30))
        assert worst <= 1e-5
        y = torch.rand(A, R, device="cuda")
        by = geo.bp(y).cpu().numpy()
        yn = y.cpu().numpy()
        for _ in range(16):
            Y, X = (int(v) for v in rng.integers(0, N, 2))
            acc_s = np.float32(0)
            acc_f = np.float32(0)
            for j in range(A):
                xmin, ymin = (Y - N // 2, X - N // 2) if j >= ns else (X - N // 2, Y - N // 2)
                _, _, v = oracle.bp_pixel_line(c[j], s[j], l[j], xmin, ymin, rz, 1.0, 2, R, yn[o[j]])
                if j >= ns:
                    acc_f = np.float32(acc_f + np.float32(v))
                else:
                    acc_s = np.float32(acc_s + np.float32(v))
            ref = np.float32(acc_s + acc_f)
            if mode == "exact":
                assert by[Y, X] == ref, (Y, X)
            assert abs(by[Y, X] - ref) <= 1e-5 * abs(ref)
        # linearity: A(2x + z) = 2Ax + Az
        z = torch.rand(N, N, device="cuda")
        lhs = geo.fp(2 * x + z).cpu().numpy()
        rhs = 2 * fx + geo.fp(z).cpu().numpy()
        assert rel_l2(lhs, rhs) < 1e-6
    finally:
        ctx.exact = False


def test_reference_entry_points(ctx):
    from synkrotomo_amd.futhark import backprojection, forwardprojection
    g = load_golden("c1.npz")
    fp = forwardprojection.forwardprojection().main(g["angles"], -45.0, 1.0, 91, g["phantom"])
    assert fp.shape == (90 * 91,)
    assert rel_l2(fp.get(), g["fp"]) <= TOL
    bp = backprojection.backprojection().main(g["angles"], -45.0, 1.0, 64, fp.get())
    assert rel_l2(bp.get(), g["bp"]) <= TOL
    with pytest.raises(ValueError):
        backprojection.backprojection().main(g["angles"], -45.0, 1.0, 64, np.zeros(91, np.float32))
    with pytest.raises(ValueError):
        forwardprojection.forwardprojection().main(g["angles"], -45.0, 1.0, 91, np.zeros(63, np.float32))
