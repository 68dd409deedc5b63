"""Measured-data preprocessing on the GPU (csrc/sk_prep.hip) against FP64 numpy statements of
the same formulas, and the rotation-axis estimate on a projected phantom."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("A,R,nf,nd", [(7, 64, 3, 2), (5, 61, 1, 0), (3, 4, 10, 10)])
def test_normalize_matches_formula(ctx, A, R, nf, nd):
    from synkrotomo_amd.preprocess import normalize
    rng = np.random.default_rng(A * R)
    flat = rng.uniform(900, 1100, (nf, R)).astype(np.float32)
    dark = rng.uniform(0, 50, (nd, R)).astype(np.float32) if nd else None
    raw = rng.uniform(-20, 1200, (2, A, R)).astype(np.float32)    # includes T <= 0 and T > 1
    dm = dark.astype(np.float64).mean(0).astype(np.float32) if nd else np.zeros(R, np.float32)
    flat[:, 1] = dm[1]                                             # dead bin: flat == dark
    raw[0, 0, 2] = np.nan
    got = normalize(raw, flat, dark, min_transmission=1e-4, scale=0.5, ctx=ctx).cpu().numpy()
    fm = flat.astype(np.float64).mean(0).astype(np.float32)
    live = fm > dm
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(live, (raw.astype(np.float64) - dm) / (fm.astype(np.float64) - dm), 0.0)
    t = np.where(np.isnan(t), 1e-4, t)
    want = -np.log(np.maximum(t, 1e-4)) / 0.5
    np.testing.assert_allclose(got[..., 1], -np.log(1e-4) / 0.5, rtol=1e-6)   # dead bin
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6)


def test_normalize_in_place_and_errors(ctx):
    from synkrotomo_amd.preprocess import normalize
    raw = torch.full((3, 8), 500.0, device="cuda")
    out = normalize(raw, torch.full((8,), 1000.0), None, out=raw, ctx=ctx)
    assert out.data_ptr() == raw.data_ptr()
    torch.testing.assert_close(raw, torch.full((3, 8), float(np.log(2.0)), device="cuda"))
    with pytest.raises(ValueError):                                # SK_EINVAL from the C ABI
        normalize(raw, torch.ones(8), None, min_transmission=0.0, ctx=ctx)
    with pytest.raises(ValueError):
        normalize(raw, torch.ones(0, 8), None, ctx=ctx)


def test_row_centroids_match_numpy(ctx):
    from synkrotomo_amd.preprocess import row_centroids
    rng = np.random.default_rng(2)
    s = rng.random((2, 9, 1000)).astype(np.float32)
    s[1, 4] = 0.0                                                 # empty row -> NaN
    got = row_centroids(s, ctx).cpu().numpy()
    r = np.arange(1000, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        want = (s.astype(np.float64) * r).sum(-1) / s.astype(np.float64).sum(-1)
    assert np.isnan(got[1, 4])
    ok = np.isfinite(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-12)


def test_rotation_centre_of_offset_scan(ctx):
    """A phantom scanned with the axis 3.3 bins left of the detector centre: the estimate
    is within 0.05 bin; the centred geometry reproduces the scan from the phantom (the
    nominal one misses it by the shift); SIRT in the centred geometry is closer to the
    phantom."""
    from synkrotomo_amd.phantom import ellipse_phantom
    from synkrotomo_amd.preprocess import centred_rhozero, rotation_centre
    from synkrotomo_amd.projector import uniform_angles
    N, A, R = 256, 180, 380
    shift = 3.3
    ang = uniform_angles(A)
    nominal = -(R - 1) / 2.0
    true_rz = nominal + shift                     # axis at bin (R-1)/2 - shift
    x = torch.from_numpy(ellipse_phantom(N, 10, seed=5)).cuda()
    sino = ctx.geometry(ang, true_rz, 1.0, R, N).fp(x)
    c0 = rotation_centre(sino, ang, ctx)
    assert abs(c0 - ((R - 1) / 2.0 - shift)) < 0.05
    data, err = {}, {}
    for name, rz in (("nominal", nominal), ("centred", centred_rhozero(c0, 1.0))):
        geo = ctx.geometry(ang, rz, 1.0, R, N)
        data[name] = rel_l2(geo.fp(x).cpu().numpy(), sino.cpu().numpy())
        img = torch.zeros(N, N, device="cuda")
        geo.sirt_(img, sino, 30)
        err[name] = rel_l2(img.cpu().numpy(), x.cpu().numpy())
    print(f"centre {c0:.4f} (true {(R - 1) / 2.0 - shift}); data misfit nominal {data['nominal']:.3e} "
          f"centred {data['centred']:.3e}; SIRT image error nominal {err['nominal']:.4f} "
          f"centred {err['centred']:.4f}")
    assert data["centred"] < 0.2 * data["nominal"]
    assert err["centred"] < err["nominal"]
