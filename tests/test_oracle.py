"""The oracle (oracle/oracle.c) pinned against what the reference and mathematics fix.

Pins, by reference function:
  sirtlib.preprocess / bp candidate selection  archive known-answer tests (+sin -> -sin remap)
  fp lengths                                   closed-form chords; brute-force exact geometry
  bp lengths (intersect_fact)                  archive KATs; adjoint identity against fp
  SIRT                                         one-step closed form, fixed point, monotone residual
  SIRT3D                                       slice k equals SIRT on slab k
"""
import math

import numpy as np
import pytest

import oracle
from conftest import load_golden, rel_l2

F32 = np.float32


def _uniform(A):
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


def _lbase(c, s, flat):
    t = F32(c / s) if flat else F32(s / c)
    return np.sqrt(F32(1) + t * t, dtype=np.float32)


def _frame(theta_a, x, y):
    """The live-code frame line for an archive (+sin convention) angle: theta_live = -theta_a."""
    th = F32(theta_a)
    c, s = F32(math.cos(float(th))), F32(math.sin(float(th)))
    flat = abs(s) >= abs(c)
    if flat:  # transposed frame, pixel coordinates swap
        return (s, c, _lbase(c, s, True), y, x)
    return (c, s, _lbase(c, s, False), x, y)


# archive/line_lib.fut:1-57 -- (pixel on 8x8, angle) -> lengths of the two candidate rays
LINE_LIB_KATS = [
    (19, 0.0, [1.0, 0.0]),
    (27, 1.57079632679, [1.0, 0.0]),
    (45, 0.78539816339, [0.17157287525, 0.65685424949]),
    (36, 0.78539816339, [1.0, 0.0]),
    (27, 0.78539816339, [1.0, 0.0]),
    (42, 2.35619449019, [0.17157287525, 0.65685424949]),
    (21, 2.35619449019, [0.65685424949, 0.17157287525]),
    (18, 0.78539816339, [0.65685424949, 0.17157287525]),
]


@pytest.mark.parametrize("pixel,angle,expected", LINE_LIB_KATS)
def test_archive_line_lib_kat(pixel, angle, expected):
    cost, sint, lb, xmin, ymin = _frame(angle, pixel % 8 - 4, pixel // 8 - 4)
    _, lengths, _ = oracle.bp_pixel_line(cost, sint, lb, xmin, ymin, -5.5, 1.0)
    np.testing.assert_allclose(lengths, expected, atol=3e-7)


# archive/projection_lib.fut:1-41 -- first candidate index angleindex*numrhos + s
PROJ_LIB_KATS = [(1, 10, 6), (3, 9, 14), (3, 6, 13), (1, 5, 5)]


@pytest.mark.parametrize("angleindex,pixel,expected", PROJ_LIB_KATS)
def test_archive_projection_lib_kat(angleindex, pixel, expected):
    angles = [0.0, 0.785398, 1.5708, 2.35619]
    cost, sint, lb, xmin, ymin = _frame(angles[angleindex], pixel % 4 - 2, pixel // 4 - 2)
    s, _, _ = oracle.bp_pixel_line(cost, sint, lb, xmin, ymin, -1.5, 1.0)
    assert angleindex * 4 + int(s[0]) == expected


def test_archive_kat_catches_sign_error():
    """A plausible mistake -- keeping the archive's +sin in the live frame -- fails a KAT."""
    th = F32(0.78539816339)
    c, s = F32(math.cos(float(th))), F32(math.sin(float(th)))
    # pixel 18 -> lowerleft (-2, -2); true lengths are [0.657, 0.172]; with -sin they differ
    _, lengths, _ = oracle.bp_pixel_line(s, F32(-c), _lbase(c, s, True), -2, -2, -5.5, 1.0)
    assert np.abs(lengths - np.array([0.65685424949, 0.17157287525])).max() > 0.1


def _chord(theta, rho, h):
    """Length of {x cos - y sin = rho} inside the square [-h, h]^2 (double)."""
    c, s = math.cos(theta), math.sin(theta)
    p0 = (rho * c, -rho * s)
    d = (s, c)
    lo, hi = -1e300, 1e300
    for k in range(2):
        if abs(d[k]) < 1e-300:
            if not (-h <= p0[k] <= h):
                return 0.0
            continue
        t1, t2 = (-h - p0[k]) / d[k], (h - p0[k]) / d[k]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    return max(0.0, hi - lo)


@pytest.mark.parametrize("N,A,R,rho0", [(64, 90, 91, -45.0), (16, 13, 25, -11.75), (32, 45, 46, -22.5)])
def test_fp_of_ones_is_the_chord(N, A, R, rho0):
    """fp(1) = length of each ray inside the image (closed form)."""
    ang = _uniform(A)
    got = oracle.fp(ang, rho0, 1.0, R, np.ones(N * N, np.float32)).reshape(A, R)
    h = N // 2
    for a in range(A):
        th = float(ang[a])
        edge_aligned = min(abs(math.sin(th)), abs(math.cos(th))) < 1e-6
        for r in range(R):
            rho = rho0 + r
            if edge_aligned and abs(abs(rho) - h) < 1e-3:
                continue  # ray on the image border: half-open pixel convention decides
            assert abs(got[a, r] - _chord(th, rho, h)) <= 2e-5 * N, (a, r)


def _exact_entry(theta, rho, px, py):
    """Exact length of the ray inside pixel [px, px+1) x [py, py+1) (double)."""
    c, s = math.cos(theta), math.sin(theta)
    p0 = (rho * c, -rho * s)
    d = (s, c)
    lo, hi = -1e300, 1e300
    for k, (a, b) in enumerate(((px, px + 1.0), (py, py + 1.0))):
        if abs(d[k]) < 1e-300:
            if not (a <= p0[k] <= b):
                return 0.0
            continue
        t1, t2 = (a - p0[k]) / d[k], (b - p0[k]) / d[k]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    return max(0.0, hi - lo)


@pytest.mark.parametrize("N,A", [(8, 7), (16, 11)])
def test_fp_matrix_matches_exact_geometry(N, A):
    """Every entry of the oracle's A equals the exact segment/pixel intersection length."""
    ang = (np.arange(A, dtype=np.float64) * np.pi / A + 0.05).astype(np.float32)
    R = int(math.ceil(math.sqrt(2) * N)) + 2
    rho0 = -(R - 1) / 2.0 + 0.25
    rowptr, cols, vals = oracle.fp_csr(ang, rho0, 1.0, R, N)
    dense = np.zeros((A * R, N * N))
    for row in range(A * R):
        for k in range(rowptr[row], rowptr[row + 1]):
            dense[row, cols[k]] += vals[k]
    h = N // 2
    exact = np.zeros_like(dense)
    for a in range(A):
        for r in range(R):
            for Y in range(N):
                for X in range(N):
                    exact[a * R + r, Y * N + X] = _exact_entry(float(ang[a]), rho0 + r, X - h, Y - h)
    assert np.abs(dense - exact).max() < 2e-5


def test_fp_matrix_catches_transposed_frame():
    """Planting a frame mistake (flat lines on the untransposed image) breaks agreement."""
    N, A = 8, 7
    ang = (np.arange(A, dtype=np.float64) * np.pi / A + 0.05).astype(np.float32)
    R, rho0 = 14, -6.25
    x = np.random.default_rng(1).random(N * N).astype(np.float32)
    good = oracle.fp(ang, rho0, 1.0, R, x)
    bad = oracle.fp(ang, rho0, 1.0, R, x.reshape(N, N).T.ravel())
    assert rel_l2(bad, good) > 1e-2


@pytest.mark.parametrize("N,A,R,rho0", [(32, 37, 47, -23.3), (24, 20, 36, -17.5)])
def test_bp_is_adjoint_of_fp(N, A, R, rho0):
    """<A x, y> = <x, B y>: the reference's bp (intersect_fact) reproduces fp's lengths.
    Geometries avoid theta = 0, pi/2 where bp drops edge rays (next test)."""
    ang = (np.arange(A, dtype=np.float64) * np.pi / A + 0.01).astype(np.float32)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(N * N).astype(np.float32)
    y = rng.standard_normal(A * R).astype(np.float32)
    lhs = np.dot(oracle.fp(ang, rho0, 1.0, R, x).astype(np.float64), y)
    rhs = np.dot(x.astype(np.float64), oracle.bp(ang, rho0, 1.0, N, y).astype(np.float64))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0)
    adj = oracle.fp_adjoint(ang, rho0, 1.0, R, N, y)
    assert rel_l2(oracle.bp(ang, rho0, 1.0, N, y), adj) < 1e-5


def test_bp_drops_edge_rays_at_theta_zero():
    """At theta = 0 with integer rho the rays lie on pixel edges; the reference's
    intersect_fact gives them length 0 (m = 0), while fp gives the full column."""
    N, R, rho0 = 16, 16, -8.0
    ang = np.array([0.0], np.float32)
    y = np.ones(R, np.float32)
    assert np.all(oracle.bp(ang, rho0, 1.0, N, y) == 0.0)
    assert np.allclose(oracle.fp_adjoint(ang, rho0, 1.0, R, N, y), 1.0)
    x = np.arange(N * N, dtype=np.float32)
    np.testing.assert_array_equal(oracle.fp(ang, rho0, 1.0, R, x), x.reshape(N, N).sum(0))


def test_sirt_one_step_closed_form():
    """One iteration from 0 is C * B(R * b) with R = 1/fp(1), C = 1/bp(1) (SIRT.fut:22-39)."""
    g = load_golden("c1.npz")
    ang, b = g["angles"], g["fp"]
    N, A, R, rz = 64, 90, 91, -45.0
    with np.errstate(divide="ignore"):
        rs = oracle.fp(ang, rz, 1.0, R, np.ones(N * N, np.float32))
        Rw = np.where(rs == 0, F32(0), F32(1) / rs).astype(np.float32)
        cs = oracle.bp(ang, rz, 1.0, N, np.ones(A * R, np.float32))
        Cw = np.where(cs == 0, F32(0), F32(1) / cs).astype(np.float32)
    q = (Rw * (b - oracle.fp(ang, rz, 1.0, R, np.zeros(N * N, np.float32)))).astype(np.float32)
    expect = (np.zeros(N * N, np.float32) + Cw * oracle.bp(ang, rz, 1.0, N, q)).astype(np.float32)
    got = oracle.sirt(ang, rz, 1.0, np.zeros(N * N, np.float32), b, 1)
    np.testing.assert_array_equal(got, expect)


def test_sirt_fixed_point_and_monotone_residual():
    g = load_golden("c1.npz")
    ang, b, x_true = g["angles"], g["fp"], g["phantom"]
    N, R, rz = 64, 91, -45.0
    x_fix = oracle.sirt(ang, rz, 1.0, x_true, b, 3)
    assert rel_l2(x_fix, x_true) < 1e-5
    with np.errstate(divide="ignore"):
        rs = oracle.fp(ang, rz, 1.0, R, np.ones(N * N, np.float32)).astype(np.float64)
        w = np.where(rs == 0, 0.0, 1.0 / rs)
    x = np.zeros(N * N, np.float32)
    prev = np.inf
    for _ in range(8):
        x = oracle.sirt(ang, rz, 1.0, x, b, 1)
        res = b - oracle.fp(ang, rz, 1.0, R, x)
        cur = float(np.sum(w * res.astype(np.float64) ** 2))
        assert cur <= prev * (1 + 1e-6)
        prev = cur


def test_sirt3d_slices_equal_sirt():
    N, A = 8, 6
    ang = _uniform(A)
    R = 12
    rhos = (np.arange(R, dtype=np.float32) - F32(5.5)).astype(np.float32)
    rng = np.random.default_rng(3)
    proj = rng.random(N * A * R).astype(np.float32)
    out = oracle.sirt3d(ang, rhos, proj, 4, N).reshape(N, N * N)
    for i in range(N):
        ref = oracle.sirt(ang, rhos[0], rhos[1] - rhos[0], np.zeros(N * N, np.float32),
                          proj[i * A * R:(i + 1) * A * R], 4)
        np.testing.assert_array_equal(out[i], ref)


def test_golden_fixtures_reproduce():
    import hashlib
    g = load_golden("fpinput512.npz")
    image = (g["image_k"].astype(np.float64) / 255.0).astype(np.float32).ravel()
    fp = oracle.fp(g["angles"], g["rhozero"], g["deltarho"], int(g["numrhos"]), image,
                   nthreads=oracle.default_threads())
    assert hashlib.sha256(fp.tobytes()).hexdigest() == str(g["fp_sha"])
    c1 = load_golden("c1.npz")
    np.testing.assert_array_equal(oracle.fp(c1["angles"], -45.0, 1.0, 91, c1["phantom"]), c1["fp"])
    np.testing.assert_array_equal(oracle.bp(c1["angles"], -45.0, 1.0, 64, c1["fp"]), c1["bp"])


def test_nnz_matches_independent_count():
    """nnz(A) for 64x64 / 90 angles / 91 bins as counted independently with numpy: 469,857."""
    rowptr, _, _ = oracle.fp_csr(_uniform(90), -45.0, 1.0, 91, 64)
    assert rowptr[-1] == 469857


def test_bp_candidates_stay_in_range_for_centred_geometry():
    for N, A in [(64, 90), (32, 33)]:
        R = int(math.ceil(math.sqrt(2) * N))
        assert oracle.bp_oob_nonzero(_uniform(A), -(R - 1) / 2.0, 1.0, R, N) == 0


def test_fp64_twin_agrees_off_edge_aligned_angles():
    """The FP64 twin matches the FP32 restatement to rounding on every angle except
    theta = pi/2: there f32 cos(pi/2) = -4.37e-8, the rays lie on pixel edges, and FP64
    keeps the true tilt while FP32 rounds it away, so the two pick neighbouring columns."""
    c1 = load_golden("c1.npz")
    f32 = c1["fp"].reshape(90, 91)
    f64 = oracle.fp_f64(c1["angles"], -45.0, 1.0, 91, c1["phantom"]).reshape(90, 91)
    keep = np.ones(90, bool)
    keep[45] = False
    assert rel_l2(f32[keep], f64[keep]) < 1e-6
    assert rel_l2(f32[45], f64[45]) > 1e-3


# ---------------------------------------------------------------- comparison tools (a14)
def test_tools_known_answers():
    """tools/sanityCheck.fut, diff.fut, mse.fut on hand-made cases, including the
    reference's quirks: diff's filter (> 0) drops index 0, and a difference of exactly the
    tolerance neither passes sanityCheck (< 0.05) nor counts in diff (> 0.05)."""
    z = np.zeros(6, np.float32)
    assert oracle.sanity_check(z, z) and oracle.diff(z, z) == 0 and oracle.mse(z, z) == 0.0
    a = z.copy()
    a[0] = 1.0                       # index 0: fails sanity, not counted by diff
    assert not oracle.sanity_check(a, z) and oracle.diff(a, z) == 0
    a[2] = np.float32(0.05)          # exactly the tolerance: fails sanity, not counted
    assert oracle.diff(a, z) == 0
    a[4] = -3.0                      # counted
    assert oracle.diff(a, z) == 1
    assert oracle.mse(a, z) == np.float32((1.0 + np.float32(0.05) ** 2 + 9.0) / 6)
    b = z.copy()
    b[3] = np.nan                    # NaN: fails sanity, not counted by diff
    assert not oracle.sanity_check(b, z) and oracle.diff(b, z) == 0
    c = z.copy()
    c[1:] = np.float32(0.049)        # inside the tolerance everywhere
    assert oracle.sanity_check(c, z) and oracle.diff(c, z) == 0


def test_tools_mse_closed_form():
    n = 1000
    d = np.arange(n, dtype=np.float32) % 7     # small integers: every square and sum exact
    assert oracle.mse(d, np.zeros(n, np.float32)) == np.float32(np.sum(d.astype(np.float64) ** 2) / n)
