"""The explicit matrices built on the GPU: index sets and values bitwise equal to the
oracle's, and SpMV results within tolerance of the dense reference outputs."""
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


GEOMS = [
    ("c1", 64, _uniform(90), 91, -45.0, 1.0),
    ("sparse5", 48, np.array([2.5, 0.1, 1.2, 0.9, 3.0], np.float32), 68, -33.5, 1.0),
    ("wide_angles", 32, np.array([-2.0, 4.0, 7.5, -0.4, 3.3, 1.5707964], np.float32), 46, -22.5, 1.0),
    ("offcentre", 32, _uniform(17), 40, 5.0, 1.0),
    ("drho_half", 24, _uniform(19), 70, -17.25, 0.5),
    ("drho_two", 40, _uniform(23), 30, -29.0, 2.0),
]


@pytest.mark.parametrize("name,N,ang,R,rz,dr", GEOMS, ids=[g[0] for g in GEOMS])
def test_csr_matches_oracle_bitwise(ctx, name, N, ang, R, rz, dr):
    geo = ctx.geometry(ang, rz, dr, R, N)
    for kind, build in (("fp", oracle.fp_csr), ("bp", oracle.bp_csr)):
        rp_ref, c_ref, v_ref = build(ang, rz, dr, R, N)
        m = geo.sparse(kind)
        assert m.nnz == int(rp_ref[-1]), kind
        rp, c, v = (t.cpu().numpy() for t in m.csr())
        np.testing.assert_array_equal(rp, rp_ref)
        np.testing.assert_array_equal(c, c_ref)
        np.testing.assert_array_equal(v, v_ref)
        assert geo.sparse_nnz(kind) == m.nnz
    geo.drop_sparse()


@pytest.mark.parametrize("name,N,ang,R,rz,dr", GEOMS, ids=[g[0] for g in GEOMS])
def test_spmv_matches_dense_reference(ctx, name, N, ang, R, rz, dr):
    rng = np.random.default_rng(12)
    x = rng.random(N * N).astype(np.float32)
    y = rng.random(ang.size * R).astype(np.float32)
    geo = ctx.geometry(ang, rz, dr, R, N)
    ref_a, ref_b = oracle.fp(ang, rz, dr, R, x), oracle.bp(ang, rz, dr, N, y)
    for variant in (0, 1, 2, 9, 10):
        geo.sparse("fp").set_spmv_variant(variant)
        ax = geo.sparse("fp").spmv(_dev(x)).cpu().numpy()
        assert rel_l2(ax, ref_a) <= TOL
        if variant == 10 and geo.ncand != 2:  # B's row slots hold two candidates per line
            continue
        geo.sparse("bp").set_spmv_variant(variant)
        by = geo.sparse("bp").spmv(_dev(y)).cpu().numpy()
        assert rel_l2(by, ref_b) <= TOL
    geo.drop_sparse()


@pytest.mark.parametrize("name,N,ang,R,rz,dr", GEOMS, ids=[g[0] for g in GEOMS])
def test_spmv_variants_bitwise(ctx, name, N, ang, R, rz, dr):
    """Every SpMV kernel that sums a row in the same order gives the same bits, batched too:
    wave per row 0 == 3 (A: steep rays gather from x^T); 16-byte streams 2 == 4 (x^T); lane
    per row 1 == 7 (SELL-64 copy) == 8 (A: SELL-64, flat rays gather from x^T) == 9 (SELL-64
    with 16-bit column steps) == 10 (row slots: A per frame row, gathers from the frame image; B
    per line, where its two candidates need deltarho in [0.71, 1.41), else ValueError on use).
    B rejects 3, 4 and 8; 5 and 6 do not exist."""
    rng = np.random.default_rng(13)
    geo = ctx.geometry(ang, rz, dr, R, N)
    for kind, ncols in (("fp", N * N), ("bp", ang.size * R)):
        x = _dev(rng.random((3, ncols)).astype(np.float32))
        M = geo.sparse(kind)
        slots = kind == "fp" or geo.ncand == 2
        variants = (0, 1, 2, 3, 4, 7, 8, 9, 10) if kind == "fp" else (0, 1, 2, 7, 9) + ((10,) if slots else ())
        out = {}
        for v in variants:
            M.set_spmv_variant(v)
            assert M.spmv_variant == v
            out[v] = M.spmv(x).cpu().numpy()
        same = [(1, 7), (1, 9)] + ([(0, 3), (2, 4), (1, 8)] if kind == "fp" else []) + ([(1, 10)] if slots else [])
        for a, b in same:
            np.testing.assert_array_equal(out[b], out[a], err_msg=f"{kind}: variant {b} vs {a}")
        if not slots:
            M.set_spmv_variant(10)
            with pytest.raises(ValueError):
                M.spmv(x)
            M.set_spmv_variant(1)
        for v in ((3, 4, 5, 6, 8, 11) if kind == "bp" else (5, 6, 11)):
            with pytest.raises(ValueError):
                M.set_spmv_variant(v)
    geo.drop_sparse()


def test_spmv_steps_beyond_16_bits(ctx):
    """Variant 9 keeps each row's columns as 16-bit steps.  In B of an 8x8 image under 40000
    bins, a pixel's entries (one or two bins per angle) lie ~40000 columns apart: asking for 9
    fails with ValueError and the autotuner leaves it out; A of the same geometry (steps of at
    most N + 1 pixels) runs it, bitwise equal to the lane-per-row kernel."""
    N, R = 8, 40000
    ang = _uniform(4)
    geo = ctx.geometry(ang, -(R - 1) / 2.0, 1.0, R, N)
    rng = np.random.default_rng(21)
    y = _dev(rng.random(ang.size * R).astype(np.float32))
    B = geo.sparse("bp")
    B.set_spmv_variant(1)
    ref = B.spmv(y).cpu().numpy()
    B.set_spmv_variant(9)
    with pytest.raises(ValueError):
        B.spmv(y)
    B.set_spmv_variant(-1)
    got = B.spmv(y).cpu().numpy()
    assert B.spmv_variant not in (-1, 9)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)  # its own summation order
    x = _dev(rng.random(N * N).astype(np.float32))
    A = geo.sparse("fp")
    A.set_spmv_variant(1)
    a1 = A.spmv(x).cpu().numpy()
    A.set_spmv_variant(9)
    np.testing.assert_array_equal(A.spmv(x).cpu().numpy(), a1)
    geo.drop_sparse()


def _sorted_entries(rows, cols, vals):
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], vals[order]


@pytest.mark.parametrize("name,N,ang,R,rz,dr", GEOMS, ids=[g[0] for g in GEOMS])
def test_adj_is_transpose_of_oracle_A_bitwise(ctx, name, N, ang, R, rz, dr):
    """SK_MAT_ADJ (A in CSC form) holds exactly the oracle A's entries, explicit zeros
    included, with bitwise values; its SpMV is the adjoint A^T.y."""
    rp_ref, c_ref, v_ref = oracle.fp_csr(ang, rz, dr, R, N)
    ray = np.repeat(np.arange(rp_ref.size - 1), np.diff(rp_ref))
    want = _sorted_entries(c_ref.astype(np.int64), ray, v_ref)   # (pixel, ray, value)
    geo = ctx.geometry(ang, rz, dr, R, N)
    m = geo.sparse("adj")
    assert (m.rows, m.cols) == (N * N, ang.size * R)
    rp, c, v = (t.cpu().numpy() for t in m.csr())
    pix = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    got = _sorted_entries(pix, c.astype(np.int64), v)
    assert m.nnz == c_ref.size
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    y = np.random.default_rng(13).random(ang.size * R).astype(np.float32)
    aty = m.spmv(_dev(y)).cpu().numpy()
    assert rel_l2(aty, oracle.fp_adjoint(ang, rz, dr, R, N, y)) <= TOL
    # lane per row on the CSR, and on the 16-bit steps over line-table order: the same bits
    m.set_spmv_variant(1)
    a1 = m.spmv(_dev(y)).cpu().numpy()
    m.set_spmv_variant(9)
    np.testing.assert_array_equal(m.spmv(_dev(y)).cpu().numpy(), a1)
    assert geo.sparse_nnz("adj") == m.nnz
    geo.drop_sparse()


def test_sparse_entry_points(ctx):
    from synkrotomo_amd.futhark import backprojection_sparse, forwardprojection_sparse
    g = load_golden("c1.npz")
    fp = forwardprojection_sparse.forwardprojection_sparse().main(g["angles"], -45.0, 1.0, 91, g["phantom"])
    assert rel_l2(fp.get(), g["fp"]) <= TOL
    bp = backprojection_sparse.backprojection_sparse().main(g["angles"], -45.0, 1.0, 64, g["fp"])
    assert rel_l2(bp.get(), g["bp"]) <= TOL


def test_batched_spmv(ctx):
    N, A, R = 32, 20, 46
    geo = ctx.geometry(_uniform(A), -22.5, 1.0, R, N)
    xs = torch.rand(3, N * N, device="cuda")
    out = geo.sparse("fp").spmv(xs)
    for i in range(3):
        torch.testing.assert_close(out[i], geo.sparse("fp").spmv(xs[i].contiguous()), rtol=0, atol=0)
    geo.drop_sparse()


def test_full_csr_512_variant_bitwise(ctx):
    """512^2, 512 angles, 725 bins (the FP64-oracle variant of the 2048^2 config):
    1.7e8 entries, every index and value equal to the oracle's."""
    N, A, R, rz = 512, 512, 725, -362.0
    ang = _uniform(A)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    rp_ref, c_ref, v_ref = oracle.fp_csr(ang, rz, 1.0, R, N)
    m = geo.sparse("fp")
    rp, c, v = (t.cpu().numpy() for t in m.csr())
    assert m.nnz == c_ref.size
    np.testing.assert_array_equal(rp, rp_ref)
    np.testing.assert_array_equal(c, c_ref)
    np.testing.assert_array_equal(v, v_ref)
    geo.drop_sparse()


def test_nnz_at_2048_matches_oracle_count(ctx):
    """nnz(A) at 2048^2 / 2048 angles / 2897 bins equals the oracle's exact count; B differs
    from A only by zero-length and theta = 0, pi/2 edge entries."""
    N, A, R, rz = 2048, 2048, 2897, -1448.0
    geo = ctx.geometry(_uniform(A), rz, 1.0, R, N)
    nnz_a = geo.sparse_nnz("fp")
    assert nnz_a == 10_937_047_089
    nnz_b = geo.sparse_nnz("bp")
    assert abs(nnz_b - nnz_a) < 1e-3 * nnz_a
