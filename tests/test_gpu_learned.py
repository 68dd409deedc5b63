"""The correction network's kernels against a CPU FP32 reference (torch conv2d on the same
BF16-rounded operands), its gradients, the GPU phantom generator, and the learned-SIRT
fallback rule."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.to(torch.bfloat16).float()


def _reference(x, net):
    """CPU FP32 forward with the kernels' rounding points: BF16 weights and activations."""
    w1 = _bf(net.w1.detach().cpu()).unsqueeze(1)                      # [32,1,3,3]
    w2 = _bf(net.w2.detach().cpu()).permute(0, 3, 1, 2)               # [32,32,3,3]
    w3 = _bf(net.w3.detach().cpu()).permute(0, 3, 1, 2)
    w4 = _bf(net.w4.detach().cpu()).permute(2, 0, 1).unsqueeze(0)     # [1,32,3,3]
    xc = x.detach().cpu().unsqueeze(1)
    h1 = _bf(F.relu(F.conv2d(xc, w1, net.b1.detach().cpu(), padding=1)))
    h2 = _bf(F.relu(F.conv2d(h1, w2, net.b2.detach().cpu(), padding=1)))
    h3 = _bf(F.relu(F.conv2d(h2, w3, net.b3.detach().cpu(), padding=1)))
    return (xc + F.conv2d(h3, w4, net.b4.detach().cpu(), padding=1)).squeeze(1)


def _net(ctx, seed=1):
    from synkrotomo_amd.learned import Correction
    net = Correction(ctx, seed=seed)
    with torch.no_grad():
        net.w4.copy_(torch.randn(net.w4.shape, generator=torch.Generator().manual_seed(9)).cuda() * 0.05)
        for b in (net.b1, net.b2, net.b3, net.b4):
            b.copy_(torch.randn(b.shape, generator=torch.Generator().manual_seed(4)).cuda() * 0.1)
    return net


@pytest.mark.parametrize("B,H,W", [(1, 16, 16), (2, 37, 23), (3, 64, 48)])
def test_forward_matches_cpu_fp32_reference(ctx, B, H, W):
    net = _net(ctx)
    x = torch.rand(B, H, W, device="cuda")
    y = net(x).detach()
    ref = _reference(x, net)
    assert rel_l2(y.cpu().numpy(), ref.numpy()) < 1e-3


def _gpu_activations(ctx, net, x):
    """The kernels' own forward activations h1, h2, h3 (BF16, NHWC) for input x [B, H, W]."""
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.projector import _ptr, _stream
    L, h, s = nat.lib(), ctx.handle, _stream(ctx.device)
    B, H, W = x.shape
    w1b, w2b, w3b = (w.detach().to(torch.bfloat16).contiguous() for w in (net.w1, net.w2, net.w3))
    h1 = torch.empty((B, H, W, 32), dtype=torch.bfloat16, device="cuda")
    h2, h3 = torch.empty_like(h1), torch.empty_like(h1)
    nat.check(L.sk_conv_in(h, B, H, W, _ptr(x), _ptr(w1b), _ptr(net.b1.detach()), _ptr(h1), s), h, "conv1")
    nat.check(L.sk_conv32(h, B, H, W, _ptr(h1), None, _ptr(w2b), _ptr(net.b2.detach()), _ptr(h2), s), h, "conv2")
    nat.check(L.sk_conv32(h, B, H, W, _ptr(h2), None, _ptr(w3b), _ptr(net.b3.detach()), _ptr(h3), s), h, "conv3")
    nchw = lambda t: t.float().permute(0, 3, 1, 2).contiguous().cpu()
    return nchw(h1), nchw(h2), nchw(h3)


def _lin_grads(conv, inp, w, b, gout):
    """FP64 gradients of a linear conv layer: (d input, d weight, d bias)."""
    ins = [t.detach().double().clone().requires_grad_(True) for t in (inp, w, b)]
    out = conv(*ins)
    return torch.autograd.grad(out, ins, gout.double())


def test_backward_matches_cpu_reference(ctx):
    """Gradients w.r.t. the input and every weight against FP64 autograd of each layer, with
    the kernels' BF16 rounding points (weights, activations, activation gradients) and the
    ReLU masks of the kernels' own forward activations (a mask recomputed on the CPU can flip
    where a pre-activation is within rounding of zero)."""
    net = _net(ctx, seed=2)
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand(2, 40, 36, device="cuda", generator=gen).requires_grad_(True)
    gy = torch.randn(2, 40, 36, device="cuda", generator=gen)
    y = net(x)
    grads = torch.autograd.grad(y, [x, net.w1, net.b1, net.w2, net.b2, net.w3, net.b3, net.w4, net.b4], gy)
    cpu = lambda t: t.detach().cpu()
    w1, w2, w3, w4 = (_bf(cpu(w)) for w in (net.w1, net.w2, net.w3, net.w4))
    b1, b2, b3, b4 = (cpu(b) for b in (net.b1, net.b2, net.b3, net.b4))
    h1, h2, h3 = _gpu_activations(ctx, net, x.detach())
    xc = cpu(x).unsqueeze(1)
    c1 = lambda inp, w, b: F.conv2d(inp, w.unsqueeze(1).contiguous(), b, padding=1)
    c32 = lambda inp, w, b: F.conv2d(inp, w.permute(0, 3, 1, 2).contiguous(), b, padding=1)
    c4 = lambda inp, w, b: F.conv2d(inp, w.permute(2, 0, 1).unsqueeze(0).contiguous(), b, padding=1)
    g = cpu(gy).unsqueeze(1)
    dh3, gw4, gb4 = _lin_grads(c4, h3, w4, b4, g)
    dp3 = _bf(dh3.float()) * (h3 > 0)               # ReLU backward with the kernel's mask
    dh2, gw3, gb3 = _lin_grads(c32, h2, w3, b3, dp3)
    dp2 = _bf(dh2.float()) * (h2 > 0)
    dh1, gw2, gb2 = _lin_grads(c32, h1, w2, b2, dp2)
    dp1 = _bf(dh1.float()) * (h1 > 0)
    dx, gw1, gb1 = _lin_grads(c1, xc, w1, b1, dp1)
    dx = dx + g.double()
    ref = [dx.squeeze(1), gw1, gb1, gw2, gb2, gw3, gb3, gw4, gb4]
    names = ["x", "w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"]
    for n, gg, r in zip(names, grads, ref):
        err = rel_l2(gg.cpu().numpy(), r.numpy())
        print(f"grad {n}: rel-L2 {err:.2e}")
        assert err < 5e-4, (n, err)


def test_two_channel_network_forward_and_backward(ctx):
    """The correction with two input channels (the image and its SIRT update direction,
    sk_conv_in2): forward against the CPU FP32 reference with the kernels' BF16 rounding points,
    and the gradients w.r.t. both inputs and the first layer against FP64 autograd (with the
    kernels' own ReLU masks)."""
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.learned import Correction
    from synkrotomo_amd.projector import _ptr, _stream
    net = Correction(ctx, seed=5, cin=2)
    with torch.no_grad():
        net.w4.copy_(torch.randn(net.w4.shape, generator=torch.Generator().manual_seed(9)).cuda() * 0.05)
        for b in (net.b1, net.b2, net.b3, net.b4):
            b.copy_(torch.randn(b.shape, generator=torch.Generator().manual_seed(4)).cuda() * 0.1)
    gen = torch.Generator(device="cuda").manual_seed(12)
    x = torch.rand(2, 40, 36, device="cuda", generator=gen).requires_grad_(True)
    d = (torch.rand(2, 40, 36, device="cuda", generator=gen) - 0.5).requires_grad_(True)
    gy = torch.randn(2, 40, 36, device="cuda", generator=gen)
    y = net(x, d)
    cpu = lambda t: t.detach().cpu()
    w1, w2, w3, w4 = (_bf(cpu(w)) for w in (net.w1, net.w2, net.w3, net.w4))
    b1, b2, b3, b4 = (cpu(b) for b in (net.b1, net.b2, net.b3, net.b4))
    inp = torch.stack([cpu(x), cpu(d)], dim=1)                          # [B, 2, H, W]
    c1 = lambda i, w, b: F.conv2d(i, w.permute(0, 3, 1, 2).contiguous(), b, padding=1)
    c32 = lambda i, w, b: F.conv2d(i, w.permute(0, 3, 1, 2).contiguous(), b, padding=1)
    c4 = lambda i, w, b: F.conv2d(i, w.permute(2, 0, 1).unsqueeze(0).contiguous(), b, padding=1)
    r1 = _bf(F.relu(c1(inp, w1, b1)))
    r2 = _bf(F.relu(c32(r1, w2, b2)))
    r3 = _bf(F.relu(c32(r2, w3, b3)))
    ref_y = (cpu(x).unsqueeze(1) + c4(r3, w4, b4)).squeeze(1)
    assert rel_l2(cpu(y).numpy(), ref_y.numpy()) < 1e-3
    grads = torch.autograd.grad(y, [x, d, net.w1, net.b1], gy)
    # the kernels' own activations (BF16 NHWC) for the masks and the layer inputs
    L, h, s = nat.lib(), ctx.handle, _stream(ctx.device)
    B, H, W = x.shape
    x2 = torch.stack([x.detach(), d.detach()], dim=-1).contiguous()
    w1b, w2b, w3b = (w.detach().to(torch.bfloat16).contiguous() for w in (net.w1, net.w2, net.w3))
    h1 = torch.empty((B, H, W, 32), dtype=torch.bfloat16, device="cuda")
    h2, h3 = torch.empty_like(h1), torch.empty_like(h1)
    nat.check(L.sk_conv_in2(h, B, H, W, _ptr(x2), _ptr(w1b), _ptr(net.b1.detach()), _ptr(h1), s), h, "conv1")
    nat.check(L.sk_conv32(h, B, H, W, _ptr(h1), None, _ptr(w2b), _ptr(net.b2.detach()), _ptr(h2), s), h, "conv2")
    nat.check(L.sk_conv32(h, B, H, W, _ptr(h2), None, _ptr(w3b), _ptr(net.b3.detach()), _ptr(h3), s), h, "conv3")
    nchw = lambda t: t.float().permute(0, 3, 1, 2).contiguous().cpu()
    h1, h2, h3 = nchw(h1), nchw(h2), nchw(h3)
    g = cpu(gy).unsqueeze(1)
    dh3, _, _ = _lin_grads(c4, h3, w4, b4, g)
    dp3 = _bf(dh3.float()) * (h3 > 0)
    dh2, _, _ = _lin_grads(c32, h2, w3, b3, dp3)
    dp2 = _bf(dh2.float()) * (h2 > 0)
    dh1, _, _ = _lin_grads(c32, h1, w2, b2, dp2)
    dp1 = _bf(dh1.float()) * (h1 > 0)
    dinp, gw1, gb1 = _lin_grads(c1, inp, w1, b1, dp1)
    ref = [dinp[:, 0] + g.double().squeeze(1), dinp[:, 1], gw1, gb1]
    for n, gg, r in zip(["x", "d", "w1", "b1"], grads, ref):
        err = rel_l2(gg.cpu().numpy(), r.numpy())
        print(f"two-channel grad {n}: rel-L2 {err:.2e}")
        assert err < 5e-4, (n, err)


def test_conv32_is_mfma_gemm_exactly_on_bf16_inputs(ctx):
    """With BF16-exact operands and sums the MFMA conv equals an FP64 convolution exactly."""
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.projector import _ptr, _stream
    B, H, W = 1, 20, 20
    x = torch.randint(-2, 3, (B, H, W, 32), device="cuda").to(torch.bfloat16)
    w = torch.randint(-1, 2, (32, 3, 3, 32), device="cuda").to(torch.bfloat16)
    y = torch.empty_like(x)
    nat.check(nat.lib().sk_conv32(ctx.handle, B, H, W, _ptr(x), None, _ptr(w), None, _ptr(y),
                                  _stream(0)), ctx.handle)
    ref = F.conv2d(x.double().cpu().permute(0, 3, 1, 2), w.double().cpu().permute(0, 3, 1, 2), padding=1)
    np.testing.assert_array_equal(y.float().cpu().permute(0, 3, 1, 2).numpy(),
                                  ref.to(torch.bfloat16).float().numpy())


def test_wgrad32_mfma_is_exact_on_small_integers(ctx):
    """The 32 -> 32 weight and bias gradients (MFMA, K = pixels) on BF16-exact small-integer
    operands, whose sums are exact in FP32 in any order: equal to the FP64 result exactly, with
    the ReLU mask applied, edge tiles on both axes (40 x 36) and two images."""
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.projector import _ptr, _stream
    B, H, W = 2, 40, 36
    gen = torch.Generator(device="cuda").manual_seed(3)
    g = torch.randint(-2, 3, (B, H, W, 32), device="cuda", generator=gen).to(torch.bfloat16)
    mk = torch.randint(-1, 2, (B, H, W, 32), device="cuda", generator=gen).to(torch.bfloat16)
    x = torch.randint(-2, 3, (B, H, W, 32), device="cuda", generator=gen).to(torch.bfloat16)
    out = torch.empty(32 * 9 * 32 + 32, device="cuda")
    nat.check(nat.lib().sk_conv_wgrad(ctx.handle, 1, B, H, W, _ptr(g), _ptr(mk), _ptr(x), _ptr(out),
                                      _stream(0)), ctx.handle)
    gm = (g.double() * (mk.double() > 0)).cpu().permute(0, 3, 1, 2)
    xc = x.double().cpu().permute(0, 3, 1, 2)
    dw = torch.nn.grad.conv2d_weight(xc, (32, 32, 3, 3), gm, padding=1)  # [co][ci][ky][kx]
    ref = torch.cat([dw.permute(0, 2, 3, 1).reshape(-1), gm.sum(dim=(0, 2, 3))])
    np.testing.assert_array_equal(out.cpu().double().numpy(), ref.numpy())


def test_gpu_phantoms_equal_numpy_generator(ctx):
    from synkrotomo_amd.learned import gpu_phantoms
    from synkrotomo_amd.phantom import ellipse_params, render
    imgs = gpu_phantoms(ctx, np.random.default_rng(77), 3, 64, 10).cpu().numpy()
    rng = np.random.default_rng(77)
    for i in range(3):
        np.testing.assert_array_equal(imgs[i], render(64, *ellipse_params(rng, 10)))


def test_learned_sirt_fallback_on_broken_network(ctx):
    """A network that returns NaN, or one that makes things worse, yields plain SIRT."""
    from synkrotomo_amd.learned import Correction, LearnedSIRT, gpu_phantoms
    from synkrotomo_amd.projector import centred_bins, uniform_angles
    N, A = 64, 45
    R, rz = centred_bins(N)
    geo = ctx.geometry(uniform_angles(A), rz, 1.0, R, N)
    x = gpu_phantoms(ctx, np.random.default_rng(5), 2, N)
    b = geo.fp(x)
    net = Correction(ctx)
    ls = LearnedSIRT(geo, net, stages=3)
    plain = torch.zeros(2, N, N, device="cuda")
    geo.sirt_(plain, b, 3)
    with torch.no_grad():
        net.b4.fill_(float("nan"))
    out, fell = ls.reconstruct(b)
    assert bool(fell.all()) or torch.allclose(out, plain, atol=1e-5)
    torch.testing.assert_close(out, plain, rtol=1e-5, atol=1e-5)
    with torch.no_grad():
        net.b4.fill_(5.0)  # a constant offset: residual far above plain SIRT's
    out, fell = ls.reconstruct(b)
    assert bool(fell.all())
    torch.testing.assert_close(out, plain, rtol=1e-5, atol=1e-5)


def test_solve_judges_by_residual_and_falls_back_to_plain(ctx):
    """LearnedSIRT.solve: every image it returns within its budget is at or below the residual
    target; with a network that spoils every stage and a budget too small to recover, each image
    falls back to plain SIRT with the same budget (bitwise: the same kernels, batch-invariant)."""
    from synkrotomo_amd.learned import Correction, LearnedSIRT, gpu_phantoms, residual
    from synkrotomo_amd.projector import centred_bins, uniform_angles
    N, A = 64, 45
    R, rz = centred_bins(N)
    geo = ctx.geometry(uniform_angles(A), rz, 1.0, R, N)
    x = gpu_phantoms(ctx, np.random.default_rng(6), 3, N)
    b = geo.fp(x)
    net = Correction(ctx, seed=3)
    with torch.no_grad():
        net.w4.zero_()
        net.b4.zero_()  # the identity correction: the stages are SIRT steps
    ls = LearnedSIRT(geo, net, stages=3)
    out, iters, fell = ls.solve(b, 0.05, 400, check_every=2)
    res = residual(geo, out, b).cpu().numpy()
    assert not bool(fell.any())
    assert np.all(res <= 0.05), res
    assert np.all((iters.cpu().numpy() >= 3) & (iters.cpu().numpy() <= 400))
    with torch.no_grad():
        # large random output weights: every stage adds high-frequency junk that 3 SIRT steps
        # cannot remove (a constant offset would not do: one SIRT step removes it)
        g = torch.Generator(device="cuda").manual_seed(9)
        net.w4.copy_(torch.randn(net.w4.shape, device="cuda", generator=g) * 100.0)
    out, iters, fell = ls.solve(b, 1e-4, 6)
    plain = torch.zeros(3, N, N, device="cuda")
    geo.sirt_(plain, b, 6)
    learned, _ = ls.reconstruct(b, guard=False)
    ls.tail_(learned, b, 3)
    assert bool((iters == 6).all())
    assert bool(fell.all()), (residual(geo, plain, b), residual(geo, learned, b))
    torch.testing.assert_close(out, plain, rtol=0, atol=0)
    with pytest.raises(ValueError):
        ls.solve(b, 0.05, 2)


def test_trained_network_gain_over_plain_sirt(ctx):
    """The shipped network against the method without it, on 16 held-out phantoms (512^2, 360
    angles, Poisson I0 = 1e4): after the 5 stages the image error is well below 5 plain SIRT
    iterations', no reconstruction falls back, and to 1.1x each phantom's noise floor the whole
    learned method (equal-cost fallback rule included) is several times faster than plain SIRT
    (5.0x median over the 1,000 test phantoms), and faster still judged by the residual."""
    from synkrotomo_amd import learned_run
    out = learned_run.evaluate(ctx, learned_run.DEFAULT_WEIGHTS, 16, batch=16, max_iters=150, verbose=False)
    after = out["after_stages"]
    tgt = out["1.1x_floor"]
    print({"after_stages": after, "speedup": tgt.get("speedup_median"),
           "equal_image_error": tgt.get("equal_image_error")})
    assert after["median_learned_err"] < 0.5 * after["median_plain_err"]
    assert out["fell_back"] == 0
    assert tgt["reached_plain"] == tgt["reached_learned"] == 16
    assert tgt["median_learned_err"] <= tgt["median_plain_err"]
    assert tgt["speedup_median"] > 3.5
    assert tgt["residual_judged"]["missed_target"] == 0
    assert tgt["residual_judged"]["speedup_median"] > tgt["speedup_median"]
    assert tgt["equal_image_error"]["speedup_median"] > 5.0
