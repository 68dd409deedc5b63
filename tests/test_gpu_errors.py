"""The C-ABI's argument checks: every bad call returns an error code with a message from
sk_last_error, before any kernel runs, and the context stays usable afterwards."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from synkrotomo_amd import _native as nat
    return nat.lib()


def _uniform(A):
    return (np.arange(A, dtype=np.float64) * np.pi / A).astype(np.float32)


@pytest.mark.parametrize("A,rz,dr,R,N,frag", [
    (0, -45.0, 1.0, 91, 64, "angle count"),
    (90, -45.0, 1.0, 0, 64, "numrhos"),
    (90, -45.0, 1.0, 91, 63, "even"),
    (90, -45.0, 1.0, 91, 0, "even"),
    (90, -45.0, 1.0, 91, 16386, "exceeds"),
    (90, -45.0, 0.0, 91, 64, "deltarho"),
    (90, -45.0, -1.0, 91, 64, "deltarho"),
    (90, 2e6, 1.0, 91, 64, "rhozero"),
    (90, -45.0, 0.05, 91, 64, "too small"),
    (90, -45.0, 1e5, 91, 64, "too large"),
])
def test_geometry_rejects(ctx, A, rz, dr, R, N, frag):
    L = _lib()
    ang = _uniform(max(A, 1))
    g = ctypes.c_void_p()
    rc = L.sk_geom_create(ctx.handle, ang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), A,
                          ctypes.c_float(rz), ctypes.c_float(dr), R, N, ctypes.byref(g))
    assert rc == 1 and not g.value
    assert frag in L.sk_last_error(ctx.handle).decode()


def test_nonfinite_angle_rejected(ctx):
    L = _lib()
    ang = _uniform(8)
    ang[3] = np.nan
    g = ctypes.c_void_p()
    rc = L.sk_geom_create(ctx.handle, ang.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 8,
                          ctypes.c_float(-11.5), ctypes.c_float(1.0), 24, 16, ctypes.byref(g))
    assert rc == 1 and "not finite" in L.sk_last_error(ctx.handle).decode()


def test_operator_calls_reject_bad_arguments(ctx):
    L = _lib()
    geo = ctx.geometry(_uniform(12), -11.5, 1.0, 24, 16)
    img = torch.zeros(16 * 16, device="cuda")
    sino = torch.zeros(12 * 24, device="cuda")
    h, gh = ctx.handle, geo.handle
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    assert L.sk_fp(h, gh, 0, p(img), p(sino), None) == 1
    assert L.sk_fp(h, gh, 1, None, p(sino), None) == 1
    assert L.sk_bp(h, gh, 70000, p(sino), p(img), None) == 1
    assert L.sk_sirt(h, gh, 1, p(img), p(sino), -1, None) == 1
    assert "iterations" in L.sk_last_error(h).decode()
    assert L.sk_geom_set_variants(gh, 8, 0) == 1
    assert L.sk_set_option(h, 99, 0) == 1
    assert L.sk_sparse_build(h, gh, 7, ctypes.byref(ctypes.c_void_p()), None) == 1
    assert L.sk_compare(h, 0, p(img), p(img), ctypes.c_float(0.05), None, None, None, None) == 1
    # the context still works
    assert L.sk_fp(h, gh, 1, p(img), p(sino), None) == 0
    torch.cuda.synchronize()
    assert float(sino.abs().sum()) == 0.0


def test_batched_rejects_too_many_groups(ctx):
    from synkrotomo_amd.projector import sirt_batched
    geo = ctx.geometry(_uniform(12), -11.5, 1.0, 24, 16)
    groups = [(geo, torch.zeros(1, 16, 16, device="cuda"), torch.zeros(1, 12, 24, device="cuda"))
              for _ in range(17)]
    with pytest.raises(ValueError):
        sirt_batched(groups, 2)
    with pytest.raises(ValueError):
        sirt_batched(groups[:1], -1)
