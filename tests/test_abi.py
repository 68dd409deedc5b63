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


def test_is_flat_tie_rule_at_f32_quarter_pi():
    """sirtlib.fut:6-7 classifies a line as flat when |sin| >= |cos|: a tie goes to the flat
    (transposed) frame.  f32(pi/4), f32(3pi/4) and their negatives are exact ties of libm's
    sinf/cosf (|sin| == |cos| == 0.70710677), and the 2048-angle benchmark set holds two of them
    (k = 512, 1536).  Both the oracle and the product's line table must put them in the flat part
    with slope exactly +-1 and line length sqrt(2); the f32 neighbours one ulp to either side
    (no longer ties) go to the side the strict comparison of the two rounded values picks.  A
    strict '>' in place of '>=' fails the first of these checks."""
    from synkrotomo_amd import _native
    m = ctypes.CDLL("libm.so.6")
    m.sinf.restype = m.cosf.restype = ctypes.c_float
    m.sinf.argtypes = m.cosf.argtypes = [ctypes.c_float]
    ties = np.array([np.pi / 4, 3 * np.pi / 4, -np.pi / 4, -3 * np.pi / 4], np.float32)
    for t in ties:
        assert abs(m.sinf(float(t))) == abs(m.cosf(float(t))), t
    bench = (np.arange(2048, dtype=np.float64) * np.pi / 2048).astype(np.float32)
    assert bench[512] == ties[0] and bench[1536] == ties[1]
    below = np.nextafter(ties, np.float32(0))          # |angle| one ulp smaller
    above = np.nextafter(ties, np.float32(np.inf) * np.sign(ties))
    angles = np.concatenate([ties, below, above]).astype(np.float32)
    A = angles.size
    c, s, l, o, ns = oracle.preprocess(angles)
    flat_orig = set(int(v) for v in o[ns:])
    lib = _native.lib()
    table = np.empty(8 * A, np.float32)
    ns_p = lib.sk_lines_compute(angles.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), A,
                                table.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    t = table.reshape(A, 8)
    flat_p = {int(t[j, 6]) for j in range(A) if int(t[j, 7]) & 1}
    assert ns_p == ns and flat_p == flat_orig
    for i in range(4):  # the ties: flat, frame (cost, sint) = (-sin, cos), slope exactly +-1
        assert i in flat_orig
        j = int(np.nonzero(o == i)[0][0])
        assert j >= ns
        assert abs(float(-s[j] / c[j])) == 1.0
        assert l[j] == np.sqrt(np.float32(2.0), dtype=np.float32)
    for i in range(4, A):  # neighbours: the classification follows the strict order of the values
        sv, cv = abs(m.sinf(float(angles[i]))), abs(m.cosf(float(angles[i])))
        assert sv != cv
        assert (i in flat_orig) == (sv > cv), (angles[i], sv, cv)
    # below pi/4 (toward 0) the line is steep, above it flat; the mirrored ties swap sides
    assert 4 not in flat_orig and 8 in flat_orig
