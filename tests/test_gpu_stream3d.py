"""Out-of-core SIRT3D: a slice stack in host memory streamed through the GPU in chunks equals
the in-core run bitwise."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chunk", [1, 5, 12, 50])
def test_streamed_equals_in_core(ctx, chunk):
    from synkrotomo_amd.projector import uniform_angles
    from synkrotomo_amd.stream3d import sirt3d_streamed
    S, N, A, R = 12, 64, 90, 91
    geo = ctx.geometry(uniform_angles(A), -45.0, 1.0, R, N)
    proj = np.random.default_rng(3).random((S, A, R)).astype(np.float32)
    x = torch.zeros(S, N, N, device="cuda")
    geo.sirt_(x, torch.from_numpy(proj).cuda(), 10)
    got = sirt3d_streamed(geo, proj, 10, chunk)
    np.testing.assert_array_equal(got, x.cpu().numpy())
    out = np.empty(S * N * N, np.float32)
    assert sirt3d_streamed(geo, proj.ravel(), 10, chunk, out=out) is out
    np.testing.assert_array_equal(out.reshape(S, N, N), got)


def test_streamed_rejects_bad_arguments(ctx):
    from synkrotomo_amd.projector import uniform_angles
    from synkrotomo_amd.stream3d import sirt3d_streamed
    geo = ctx.geometry(uniform_angles(10), -7.5, 1.0, 16, 10)
    with pytest.raises(ValueError):
        sirt3d_streamed(geo, np.zeros(10 * 16 + 1, np.float32), 2, 4)
    with pytest.raises(TypeError):
        sirt3d_streamed(geo, np.zeros(10 * 16, np.float64), 2, 4)
    with pytest.raises(ValueError):
        sirt3d_streamed(geo, np.zeros(10 * 16, np.float32), 2, 0)
