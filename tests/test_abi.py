"""The C-ABI library: it loads without a GPU, exports every declared symbol, and its host
geometry equals the oracle's preprocess (sirtlib.fut:28-44) bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "synkro.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from synkrotomo_amd import _native
    lib = _native.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.SIGNATURES), "ctypes signatures out of sync with synkro.h"


@pytest.mark.parametrize("angles", [
    (np.arange(90, dtype=np.float64) * np.pi / 90).astype(np.float32),
    (np.arange(2048, dtype=np.float64) * np.pi / 2048).astype(np.float32),
    (np.pi / 4 + np.arange(805, dtype=np.float64) * (np.pi / 2 / 805)).astype(np.float32),
    np.random.default_rng(5).uniform(-7, 7, 257).astype(np.float32),
])
def test_host_line_table_matches_oracle(angles):
    from synkrotomo_amd import _native
    lib = _native.lib()
    A = angles.size
    table = np.empty(8 * A, np.float32)
    ns = lib.sk_lines_compute(angles.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), A,
                              table.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    c, s, l, o, ns_ref = oracle.preprocess(angles)
    t = table.reshape(A, 8)
    assert ns == ns_ref
    np.testing.assert_array_equal(t[:, 0], c)
    np.testing.assert_array_equal(t[:, 1], s)
    np.testing.assert_array_equal(t[:, 2], l)
    np.testing.assert_array_equal(t[:, 6].astype(np.int32), o)
    np.testing.assert_array_equal(t[:, 3], (-s / c).astype(np.float32))


def test_entry_points_reject_wrong_types_before_touching_the_gpu():
    from synkrotomo_amd.futhark import SIRT, backprojection, forwardprojection
    ang64 = np.zeros(4, np.float64)
    with pytest.raises(TypeError):
        forwardprojection.forwardprojection().main(ang64, -1.5, 1.0, 4, np.zeros(16, np.float32))
    with pytest.raises(TypeError):
        forwardprojection.forwardprojection().main(np.zeros(4, np.float32), "x", 1.0, 4,
                                                   np.zeros(16, np.float32))
    with pytest.raises(TypeError):
        backprojection.backprojection().main(np.zeros(4, np.float32), -1.5, 1.0, 4.5,
                                             np.zeros(16, np.float32))
    with pytest.raises(TypeError):
        SIRT.SIRT().main(np.zeros(4, np.float32), -1.5, 1.0, np.zeros((4, 4), np.float32),
                         np.zeros(16, np.float32), 3)
