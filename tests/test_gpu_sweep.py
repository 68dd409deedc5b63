"""A small size sweep on the GPU writes the reference's result-file format and compares."""
import json
import os

import pytest

from conftest import GOLDEN
from synkrotomo_amd import sweep

pytestmark = pytest.mark.gpu


def test_sweep_writes_results_and_speedups(tmp_path, capsys):
    ref = os.path.join(GOLDEN, "ref_results_rtx2080ti")
    assert sweep.main(["--sizes", "64,128", "--runs", "2", "--angles", "full", "--out", str(tmp_path),
                       "--compare", ref]) == 0
    rec = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    fp = sweep.read_results(str(tmp_path / "fp.csv"))
    bp = sweep.read_results(str(tmp_path / "bp.csv"))
    assert sorted(fp) == sorted(bp) == [64, 128] and all(v[0] > 0 for v in fp.values())
    assert rec["gups"]["128"]["angles"] == 202 and rec["gups"]["128"]["bins"] == 182
    for k in ("fut_fp", "fut_bp", "astra_fp", "astra_bp"):
        assert set(rec["speedup_vs_reference_files"][k]) == {"64", "128"}
