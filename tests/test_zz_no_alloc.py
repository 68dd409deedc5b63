"""sk_fp_noalloc: the forward projector with no device allocation, called through the C ABI.

Its results must be bitwise those of sk_fp (the same kernels; only where the transposed image
of the flat lines is staged differs), on two calls in a row with the second smaller, and the
call must not change the device's free memory.  Inputs beyond its fixed scratch are rejected
before any work is enqueued.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _uniform(a):
    return (np.arange(a, dtype=np.float64) * np.pi / a).astype(np.float32)


def _noalloc(geo, img, out):
    from synkrotomo_amd import _native as nat
    from synkrotomo_amd.projector import _ptr, _stream
    batch = img.numel() // (geo.N * geo.N)
    return nat.lib().sk_fp_noalloc(geo.ctx.handle, geo.handle, batch, _ptr(img), _ptr(out),
                                   _stream(geo.ctx.device))


def _err(geo):
    from synkrotomo_amd import _native as nat
    return nat.lib().sk_last_error(geo.ctx.handle).decode()


@pytest.mark.parametrize("fp_variant", [-1, 7], ids=["default", "grouped"])
def test_fp_noalloc_equals_fp_twice_second_smaller(fp_variant):
    from synkrotomo_amd.projector import Context
    ctx = Context(0)  # a fresh context: its own scratch has never been grown
    try:
        gen = torch.Generator(device="cuda").manual_seed(77)
        cases = [(512, 360, 725, -362.0, 4), (128, 90, 182, -90.5, 1)]  # second call smaller
        prepared = []
        for N, A, R, rz, b in cases:
            geo = ctx.geometry(_uniform(A), rz, 1.0, R, N)
            geo.set_variants(fp_variant, -1)
            x = torch.rand(b, N, N, device="cuda", generator=gen)
            prepared.append((geo, x, torch.full((b, A, R), float("nan"), device="cuda")))
        # The HIP runtime sets up per-kernel state (argument buffers and the like) on a kernel's
        # first launch in the process; one tiny call per geometry first, so that the free-memory
        # check below sees only what the calls themselves would allocate.
        for geo, x, out in prepared:
            warm = torch.empty((1, geo.A, geo.R), device="cuda")
            assert _noalloc(geo, x[:1].contiguous(), warm) == 0, _err(geo)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for geo, x, out in prepared:
            print(f"sk_fp_noalloc N={geo.N} A={geo.A} batch={x.shape[0]}", flush=True)
            assert _noalloc(geo, x, out) == 0, _err(geo)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        for geo, x, out in prepared:
            ref = geo.fp(x)  # sk_fp: the context scratch (allocated now) holds x^T
            torch.testing.assert_close(out, ref, rtol=0, atol=0)
            ref0 = oracle.fp(geo.angles, geo.rho0, geo.drho, geo.R, x[0].cpu().numpy())
            assert rel_l2(out[0].cpu().numpy(), ref0) <= 1e-5
        # the same two calls again, after sk_fp ran: still bitwise
        for geo, x, out in prepared:
            out2 = torch.empty_like(out)
            assert _noalloc(geo, x, out2) == 0, _err(geo)
            torch.testing.assert_close(out2, out, rtol=0, atol=0)
    finally:
        ctx.close()


def test_fp_noalloc_rejects_inputs_beyond_its_scratch(ctx):
    from synkrotomo_amd import _native as nat
    N, A, R = 2048, 16, 2897
    geo = ctx.geometry(_uniform(A), -1448.0, 1.0, R, N)  # holds flat angles
    assert geo.n_steep < A
    x = torch.zeros(5, N, N, device="cuda")  # 5 * 2048^2 floats > 16 Mi
    out = torch.full((5, A, R), 7.0, device="cuda")
    rc = _noalloc(geo, x, out)
    assert rc == nat.SK_EINVAL
    assert "exceeds its fixed scratch" in _err(geo)
    torch.cuda.synchronize()
    assert bool(torch.all(out == 7.0))  # nothing was launched
    # the largest accepted batch at this size runs
    x4, out4 = x[:4].contiguous(), torch.empty(4, A, R, device="cuda")
    assert _noalloc(geo, x4, out4) == 0, _err(geo)
    torch.testing.assert_close(out4, geo.fp(x4), rtol=0, atol=0)
