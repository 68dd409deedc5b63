"""Datasets in Futhark's value formats run through the drop-in entry points
(synkrotomo_amd.futhark_data: the `futhark run` / `futhark bench` of this package)."""
import json

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from synkrotomo_amd import futhark_data as fd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binary", [True, False], ids=["binary", "text"])
def test_dataset_through_entry_points(tmp_path, binary):
    """The 64^2 / 90-angle configuration written as datasets (arguments in entry-point order):
    forwardprojection, backprojection and SIRT on them equal the oracle's outputs."""
    g = load_golden("c1.npz")
    ang, rz, dr, R, N = g["angles"], np.float32(-45.0), np.float32(1.0), np.int32(91), np.int32(64)
    cases = [("forwardprojection", [ang, rz, dr, R, g["phantom"]], g["fp"]),
             ("backprojection", [ang, rz, dr, N, g["fp"]], g["bp"]),
             ("SIRT", [ang, rz, dr, np.zeros(64 * 64, np.float32), g["fp"], np.int32(10)], g["sirt10"])]
    for name, args, want in cases:
        p = tmp_path / f"{name}.in"
        fd.dump(args, str(p), binary=binary)
        got = fd.run(name, fd.load(str(p)))
        assert got.dtype == np.float32 and got.shape == want.shape
        assert rel_l2(got, want) <= 1e-5, name


def test_cli_run_and_bench(tmp_path, capsys):
    g = load_golden("c1.npz")
    p = tmp_path / "fpinput"
    fd.dump([g["angles"], np.float32(-45.0), np.float32(1.0), np.int32(91), g["phantom"]], str(p),
            binary=False)
    out = tmp_path / "fp.out"
    assert fd._cli(["run", "forwardprojection", str(p), "-o", str(out), "--binary"]) == 0
    back = fd.load(str(out))
    assert len(back) == 1 and rel_l2(back[0], g["fp"]) <= 1e-5
    assert fd._cli(["bench", "forwardprojection", str(p), "--runs", "3"]) == 0
    rec = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert rec["entry"] == "forwardprojection" and rec["runs"] == 3 and rec["mean_us"] > 0
