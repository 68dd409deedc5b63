"""Host side of the measured-data preprocessing: the rotation-axis fit."""
import numpy as np
import pytest

from synkrotomo_amd.preprocess import centred_rhozero, fit_centre


def test_fit_centre_recovers_sinusoid_offset():
    rng = np.random.default_rng(4)
    th = np.arange(180) * np.pi / 180
    c = 361.7 + 40.0 * np.cos(th) - 25.0 * np.sin(th) + rng.normal(0, 0.01, th.size)
    c[[3, 50]] = np.nan                       # empty projections are left out
    assert abs(fit_centre(c, th) - 361.7) < 0.01


def test_fit_centre_needs_three_projections():
    with pytest.raises(ValueError):
        fit_centre([1.0, np.nan, 2.0], [0.0, 0.5, 1.0])
    with pytest.raises(ValueError):
        fit_centre([1.0, 2.0], [0.0])


def test_centred_rhozero():
    assert centred_rhozero(362.0, 1.0) == -362.0
    assert centred_rhozero(10.0, 0.5) == -5.0
