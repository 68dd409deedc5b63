"""Futhark value formats (synkrotomo_amd.futhark_data): the published binary layout byte for
byte, the text grammar, round trips of every element type, and the reference's own dataset."""
import os
import struct

import numpy as np
import pytest

from conftest import load_golden
from synkrotomo_amd import futhark_data as fd

REF_DATASET = "/root/reference/futhark/data/fpinputf32rad512"


def test_binary_layout_known_answer():
    """'b', version 2, rank, 4-byte type name, int64 shape, little-endian data."""
    a = np.array([[1.0, -0.5, 3.25], [0.0, 2.0, -8.0]], np.float32)
    want = b"b\x02\x02 f32" + struct.pack("<2q", 2, 3) + struct.pack("<6f", *a.ravel())
    assert fd.dumps([a]) == want
    assert fd.dumps([np.int32(725)]) == b"b\x02\x00 i32" + struct.pack("<i", 725)
    assert fd.dumps([np.array([True, False])]) == b"b\x02\x01bool" + struct.pack("<q", 2) + b"\x01\x00"
    assert fd.dumps([np.array([7], np.uint8)]) == b"b\x02\x01  u8" + struct.pack("<q", 1) + b"\x07"
    v = fd.loads(want)
    assert len(v) == 1 and v[0].dtype == np.float32 and v[0].shape == (2, 3)
    np.testing.assert_array_equal(v[0], a)


def test_text_grammar():
    src = """-- a dataset in the reference's style
    [0.7853981633974483f32,0.7873494631822867f32]
    -362.0f32 1.0f32 725i32
    3 3.5 true false f32.nan -f64.inf 0x1Fi64 1_000u16 2f32
    [[1i8, -2i8], [3i8, 4i8]] empty([0][3]f32) [1.5, -2.5]
    """
    v = fd.loads(src)
    assert v[0].dtype == np.float32 and v[0].shape == (2,)
    assert v[0][0] == np.float32(0.7853981633974483)
    assert v[1] == np.float32(-362.0) and v[1].dtype == np.float32
    assert v[3] == 725 and v[3].dtype == np.int32
    assert v[4].dtype == np.int32 and v[5].dtype == np.float64 and v[5] == 3.5
    assert v[6] == np.True_ and v[7] == np.False_
    assert np.isnan(v[8]) and v[8].dtype == np.float32
    assert v[9] == -np.inf and v[9].dtype == np.float64
    assert v[10] == 31 and v[10].dtype == np.int64
    assert v[11] == 1000 and v[11].dtype == np.uint16
    assert v[12] == 2.0 and v[12].dtype == np.float32
    np.testing.assert_array_equal(v[13], np.array([[1, -2], [3, 4]], np.int8))
    assert v[14].shape == (0, 3) and v[14].dtype == np.float32
    assert v[15].dtype == np.float64
    assert len(v) == 16


@pytest.mark.parametrize("bad", ["[1i32, 2f32]", "[[1, 2], [3]]", "[1, [2]]", "[]", "[1, 2",
                                 "1.5i32", "300i8", "empty([2]f32)", "[1 2]"])
def test_text_errors(bad):
    with pytest.raises((ValueError, OverflowError)):
        fd.loads(bad)


def test_binary_errors():
    good = fd.dumps([np.arange(4, dtype=np.int32)])
    with pytest.raises(ValueError):
        fd.loads(good[:-1])                       # truncated data
    with pytest.raises(ValueError):
        fd.loads(good[:3] + b" i31" + good[7:])   # unknown element type
    with pytest.raises(ValueError):
        fd.loads(b"b\x02\x01 f32" + struct.pack("<q", -1))


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip_every_type(binary):
    rng = np.random.default_rng(3)
    vals = []
    for t, dt in fd._TYPES.items():
        if dt.kind == "f":
            a = rng.standard_normal((3, 4)).astype(dt) * dt.type(1e3)
            a[0, 0], a[0, 1], a[0, 2] = np.nan, np.inf, -np.inf
        elif dt.kind == "b":
            a = rng.random((3, 4)) < 0.5
        else:
            info = np.iinfo(dt)
            a = rng.integers(info.min, info.max, size=(3, 4), dtype=dt, endpoint=True)
        vals += [a, a[0], a[1, 1], np.zeros((0, 2), dt)]
    back = fd.loads(fd.dumps(vals, binary=binary))
    assert len(back) == len(vals)
    for x, y in zip(vals, back):
        x = np.asarray(x)
        assert y.dtype == x.dtype and np.shape(y) == x.shape
        np.testing.assert_array_equal(np.asarray(y), x)   # NaN == NaN here


def test_mixed_text_and_binary_stream(tmp_path):
    a = np.linspace(-1, 1, 7, dtype=np.float32)
    data = fd.dumps([a], binary=False).encode() + fd.dumps([np.int32(5)]) + b"\n 2.5f32 " + fd.dumps([a])
    p = tmp_path / "mixed"
    p.write_bytes(data)
    v = fd.load(str(p))
    assert len(v) == 4 and v[1] == 5 and v[2] == np.float32(2.5)
    np.testing.assert_array_equal(v[0], a)
    np.testing.assert_array_equal(v[3], a)


def test_text_floats_round_trip_bitwise():
    """Every float32 bit pattern class (subnormal, max, tiny steps) survives text."""
    bits = np.array([1, 2, 0x007FFFFF, 0x00800000, 0x7F7FFFFF, 0x3F800001, 0x3EAAAAAB, 0xBF7FFFFF],
                    np.uint32)
    a = bits.view(np.float32)
    b = fd.loads(fd.dumps([a], binary=False))[0]
    np.testing.assert_array_equal(b.view(np.uint32), bits)


@pytest.mark.skipif(not os.path.exists(REF_DATASET), reason="reference tree not mounted")
def test_reference_dataset_matches_fixture():
    """The reference's own benchmark input parses to the committed fixture's values."""
    v = fd.load(REF_DATASET)
    g = load_golden("fpinput512.npz")
    assert [np.asarray(x).dtype for x in v] == [np.float32, np.float32, np.float32, np.int32, np.float32]
    np.testing.assert_array_equal(v[0], g["angles"])
    assert v[1] == g["rhozero"] and v[2] == g["deltarho"] and v[3] == g["numrhos"]
    img = (g["image_k"].astype(np.float64) / 255.0).astype(np.float32).ravel()
    np.testing.assert_array_equal(v[4], img)


def test_entry_point_names():
    assert set(fd.ENTRIES) == {"forwardprojection", "backprojection", "forwardprojection_sparse",
                               "backprojection_sparse", "SIRT", "SIRT3D"}
    with pytest.raises(ValueError):
        fd.entry_point("nope")
