"""Size sweeps in the reference's result-file format (synkrotomo_amd.sweep): the dataset rule,
the size;us files and the speedup against the reference's own results."""
import os

import numpy as np

from conftest import GOLDEN, load_golden
from synkrotomo_amd import sweep


def test_dataset_rule_reproduces_reference_512():
    """The reference's fpinputf32rad512 (805 angles in [pi/4, 3pi/4), 725 bins, rhozero -362)
    is the sweep's geometry for N = 512, angle for angle."""
    g = load_golden("fpinput512.npz")
    th, rz, dr, R = sweep.sweep_geometry(512)
    np.testing.assert_array_equal(th, g["angles"])
    assert (rz, dr, R) == (float(g["rhozero"]), float(g["deltarho"]), int(g["numrhos"]))
    th2, _, _, R2 = sweep.sweep_geometry(2048, "full")
    assert th2.size == 3217 and R2 == 2897 and th2[0] == 0 and th2[-1] < np.float32(np.pi)


def test_results_files_round_trip_and_speedup(tmp_path):
    p = tmp_path / "fp.csv"
    sweep.write_results(str(p), [(64, 12.345), (128, 100.0)])
    assert p.read_text() == "64;12.35\n128;100.00\n"
    assert sweep.read_results(str(p)) == {64: [12.35], 128: [100.0]}
    ref = sweep.read_results(os.path.join(GOLDEN, "ref_results_rtx2080ti", "fut_fp.csv"))
    assert sorted(ref) == list(sweep.REFERENCE_SIZES) and all(len(v) == 10 for v in ref.values())
    s = sweep.speedups({2048: 1000.0, 7: 1.0}, ref)
    assert list(s) == [2048] and s[2048] == round(float(np.median(ref[2048])) / 1000.0, 2)
