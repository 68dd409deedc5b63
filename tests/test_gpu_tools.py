"""The reference's comparison tools (futhark/tools/*.fut) on the GPU (sk_compare) against the
oracle's sequential restatement."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(21)
    n = 64 * 64
    sane = rng.random(n).astype(np.float32)
    same = sane.copy()
    near = sane + rng.uniform(-0.049, 0.049, n).astype(np.float32)
    far = near.copy()
    idx = rng.choice(np.arange(1, n), 37, replace=False)
    far[idx] += np.float32(0.5)
    far[0] += np.float32(1.0)                   # index 0: dropped by diff's filter
    edge = sane.copy()
    edge[5] = sane[5] + np.float32(0.05)        # about the tolerance; the oracle decides
    nan = sane.copy()
    nan[17] = np.nan
    return [("same", sane, same), ("near", sane, near), ("far", sane, far),
            ("edge", sane, edge), ("nan", sane, nan)]


@pytest.mark.parametrize("name,sane,test", _cases(), ids=[c[0] for c in _cases()])
def test_tools_match_oracle(name, sane, test):
    from synkrotomo_amd.futhark.tools import diff, mse, sanityCheck
    assert bool(sanityCheck.sanityCheck().main(sane, test)) == oracle.sanity_check(sane, test)
    assert int(diff.diff().main(sane, test)) == oracle.diff(sane, test)
    m = float(mse.mse().main(sane, test))
    ref = oracle.mse(sane, test)
    if np.isnan(ref):
        assert np.isnan(m)
    else:
        assert abs(m - ref) <= 1e-5 * max(abs(ref), 1e-30)
        d = sane.astype(np.float64) - test.astype(np.float64)
        assert abs(m - np.mean(d * d)) <= 1e-6 * max(np.mean(d * d), 1e-30)


def test_compare_large_and_device_inputs():
    """4M entries (a 2048^2 image) given as device tensors; mse against an FP64 mean."""
    from synkrotomo_amd.compare import compare
    g = torch.Generator(device="cuda").manual_seed(3)
    a = torch.rand(2048 * 2048, device="cuda", generator=g)
    b = a + (torch.rand(a.shape, device="cuda", generator=g) - 0.5) * 0.2
    ok, over, m = compare(a, b)
    d = (a.double() - b.double())
    assert ok == bool((d.abs().float() < 0.05).all())
    idx = torch.nonzero((a - b).abs() > 0.05).flatten()
    assert over == int((idx > 0).sum())
    ref = float(((a - b) * (a - b)).double().mean())
    assert abs(m - ref) <= 1e-6 * ref


def test_compare_argument_errors():
    from synkrotomo_amd.compare import compare
    from synkrotomo_amd.futhark.tools import mse
    a = torch.zeros(8, device="cuda")
    with pytest.raises(ValueError):
        compare(a, torch.zeros(9, device="cuda"))
    with pytest.raises(TypeError):
        compare(a, torch.zeros(8, device="cuda", dtype=torch.float64))
    with pytest.raises(TypeError):
        mse.mse().main(np.zeros(4, np.float64), np.zeros(4, np.float32))
