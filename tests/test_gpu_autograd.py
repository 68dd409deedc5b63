"""Differentiable projector and SIRT steps: exact adjoints, FP64 finite differences,
agreement with the reference backprojection, checkpointed unrolled SIRT."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _uniform(A, offset=0.0):
    return (np.arange(A, dtype=np.float64) * np.pi / A + offset).astype(np.float32)


def _dot(a, b):
    return float(torch.sum(a.double() * b.double()))


def test_adjoint_dot_product_batch16_512(ctx):
    """Differentiable-projector config: 16 x 512^2, 360 angles, 725 bins, x, y ~ N(0, 1)."""
    from synkrotomo_amd.autograd import project
    N, A, R = 512, 360, 725
    geo = ctx.geometry(_uniform(A), -362.0, 1.0, R, N)
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(16, N, N, device="cuda", generator=gen, requires_grad=True)
    y = torch.randn(16, A, R, device="cuda", generator=gen)
    ax = project(x, geo)
    (gx,) = torch.autograd.grad(ax, x, y)
    lhs, rhs = _dot(ax, y), _dot(x, gx)
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs)), (lhs, rhs)


def test_backproject_adjoint_and_oracle_transpose(ctx):
    """B^T (sk_bp_transpose) against <B y, v> = <y, B^T v> and the oracle's CSR of B."""
    from synkrotomo_amd.autograd import backproject
    N, A, R, rz = 64, 90, 91, -45.0   # includes theta = 0, pi/2
    ang = _uniform(A)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    rng = np.random.default_rng(3)
    y = torch.from_numpy(rng.standard_normal((A, R)).astype(np.float32)).cuda().requires_grad_(True)
    v = torch.from_numpy(rng.standard_normal((N, N)).astype(np.float32)).cuda()
    by = backproject(y, geo)
    (gy,) = torch.autograd.grad(by, y, v)
    lhs, rhs = _dot(by, v), _dot(y, gy)
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))
    rp, c, vals = oracle.bp_csr(ang, rz, 1.0, R, N)
    B = sp.csr_matrix((vals.astype(np.float64), c, rp), shape=(N * N, A * R))
    ref = B.T @ v.cpu().numpy().ravel().astype(np.float64)
    assert rel_l2(gy.cpu().numpy().ravel(), ref) <= 1e-5


@pytest.mark.parametrize("offset", [0.0, 0.01])
def test_fp64_central_differences(ctx, offset):
    """d<A x, w>/dx by FP64 central differences (h = 1e-4) on the FP64 twin of the forward
    projector, 64^2, 90 angles, 40 sampled pixels.  On the uniform angle set the theta = pi/2
    row of w is zero: there the FP64 twin resolves the edge-aligned rays into the
    neighbouring column (f32 rounds the tilt away), so the FP64 operator itself differs."""
    from synkrotomo_amd.autograd import project
    N, A, R, rz = 64, 90, 91, -45.0
    ang = _uniform(A, offset)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    rng = np.random.default_rng(5)
    x0 = rng.random(N * N)
    w = rng.standard_normal((A, R))
    if offset == 0.0:
        w[45] = 0.0
    xt = torch.from_numpy(x0.astype(np.float32)).cuda().reshape(N, N).requires_grad_(True)
    (g,) = torch.autograd.grad(project(xt, geo), xt, torch.from_numpy(w.astype(np.float32)).cuda())
    g = g.cpu().numpy().ravel()
    h = 1e-4
    for i in rng.choice(N * N, 40, replace=False):
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fd = (np.dot(oracle.fp_f64(ang, rz, 1.0, R, xp), w.ravel()) -
              np.dot(oracle.fp_f64(ang, rz, 1.0, R, xm), w.ravel())) / (2 * h)
        assert abs(g[i] - fd) <= 1e-4 * (1.0 + abs(fd)), (i, g[i], fd)


def test_backward_against_reference_backprojection(ctx):
    """A^T equals the reference's backprojection except on theta = 0, pi/2."""
    from synkrotomo_amd.autograd import project
    N, A, R, rz = 64, 90, 91, -45.0
    ang = _uniform(A)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    y = np.random.default_rng(8).random((A, R)).astype(np.float32)
    y[0] = 0.0
    y[45] = 0.0
    x = torch.zeros(N, N, device="cuda", requires_grad=True)
    (g,) = torch.autograd.grad(project(x, geo), x, torch.from_numpy(y).cuda())
    assert rel_l2(g.cpu().numpy().ravel(), oracle.bp(ang, rz, 1.0, N, y)) <= 1e-5


def test_sirt_step_matches_forward_and_oracle(ctx):
    from synkrotomo_amd.autograd import SirtWeights, sirt_step
    N, A, R, rz = 64, 90, 91, -45.0
    ang = _uniform(A)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    b = np.random.default_rng(1).random((A, R)).astype(np.float32)
    w = SirtWeights(geo)
    x = torch.zeros(N, N, device="cuda")
    bt = torch.from_numpy(b).cuda()
    for _ in range(4):
        x = sirt_step(x, bt, geo, w)
    ref = oracle.sirt(ang, rz, 1.0, np.zeros(N * N, np.float32), b, 4)
    assert rel_l2(x.cpu().numpy().ravel(), ref) <= 1e-5


def test_unrolled_sirt_gradient_fd_fp64(ctx):
    """10 unrolled steps: gradient w.r.t. x0 and b against FP64 central differences of the
    oracle's FP64 SIRT (offset angles: no edge-aligned rays)."""
    from synkrotomo_amd.autograd import unrolled_sirt
    N, A, R, rz = 32, 45, 46, -22.5
    ang = _uniform(A, 0.013)
    geo = ctx.geometry(ang, rz, 1.0, R, N)
    rng = np.random.default_rng(21)
    x0 = rng.random(N * N) * 0.1
    b = rng.random(A * R)
    wout = rng.standard_normal(N * N)
    xt = torch.from_numpy(x0.astype(np.float32)).cuda().reshape(N, N).requires_grad_(True)
    bt = torch.from_numpy(b.astype(np.float32)).cuda().reshape(A, R).requires_grad_(True)
    out = unrolled_sirt(xt, bt, geo, 10)
    gx, gb = torch.autograd.grad(out, (xt, bt), torch.from_numpy(wout.astype(np.float32)).cuda().reshape(N, N))
    gx, gb = gx.cpu().numpy().ravel(), gb.cpu().numpy().ravel()

    def L(xx, bb):
        return float(np.dot(oracle.sirt_f64(ang, rz, 1.0, xx, bb, 10), wout))

    h = 1e-4
    for i in rng.choice(N * N, 12, replace=False):
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fd = (L(xp, b) - L(xm, b)) / (2 * h)
        assert abs(gx[i] - fd) <= 1e-3 * (1.0 + abs(fd)), ("x", i, gx[i], fd)
    for i in rng.choice(A * R, 12, replace=False):
        bp_, bm = b.copy(), b.copy()
        bp_[i] += h
        bm[i] -= h
        fd = (L(x0, bp_) - L(x0, bm)) / (2 * h)
        assert abs(gb[i] - fd) <= 1e-3 * (1.0 + abs(fd)), ("b", i, gb[i], fd)


def test_unrolled_sirt_checkpointing_gives_same_gradients(ctx):
    """C4 shape: 16 x 512^2, 360 angles, 10 unrolled steps with a per-step correction that
    has 8x image-sized intermediates; checkpointed and plain backward agree, and the
    checkpointed one keeps less activation memory."""
    from synkrotomo_amd.autograd import unrolled_sirt
    N, A, R = 512, 360, 725
    geo = ctx.geometry(_uniform(A), -362.0, 1.0, R, N)
    gen = torch.Generator(device="cuda").manual_seed(3)
    b = torch.rand(16, A, R, device="cuda", generator=gen)
    scales = torch.linspace(0.5, 1.5, 8, device="cuda").view(1, 8, 1, 1)

    def corr(x):
        return x + 0.01 * torch.tanh(x.unsqueeze(1) * scales).mean(1)

    grads = []
    peaks = []
    for ck in (False, True):
        x0 = torch.zeros(16, N, N, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        out = unrolled_sirt(x0, b, geo, 10, correction=corr, checkpoint_segments=ck)
        loss = out.square().mean()
        loss.backward()
        torch.cuda.synchronize()
        peaks.append(torch.cuda.max_memory_allocated())
        grads.append(x0.grad.clone())
    assert rel_l2(grads[0].cpu().numpy(), grads[1].cpu().numpy()) <= 1e-6
    assert peaks[1] < 0.5 * peaks[0], peaks
