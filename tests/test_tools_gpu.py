"""The reference-protocol CLI on the GPU, in both threshold modes, and the
table generator built on it (the harness of generate_compare_tables.py)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mst-artifact-sc23_amd")
sys.path.insert(0, PKG)

pytestmark = pytest.mark.gpu


def test_cli_modes_and_tables(tmp_path):
    import tables
    egr = tmp_path / "rmat12.egr"
    r = subprocess.run([os.path.join(PKG, "bin", "ecl_gen"), "rmat", "12", "16", "0.57", "0.19", "0.19", "2", "1",
                        str(1 << 24), str(egr)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    grid = tmp_path / "grid.egr"
    r = subprocess.run([os.path.join(PKG, "bin", "ecl_gen"), "grid", "64", "64", "1", "1", str(1 << 20), str(grid)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("step: tables.run", flush=True)
    runtime, thru = tables.run(str(tmp_path), [str(egr), str(grid)])
    for col in ("ecl_mst", "ecl_mst_learned"):
        cap = (tmp_path / f"{col}_out.csv").read_text()
        assert "ERROR" not in cap and cap.count("input:") == 2
        for p in (str(egr), str(grid)):
            assert runtime[col][p] > 0 and thru[col][p] > 0
    lines = (tmp_path / "comparison_throughputs.csv").read_text().splitlines()
    assert lines[0] == ",ecl_mst, ecl_mst_learned, "
    assert [ln.split(",")[0] for ln in lines[1:]] == ["rmat12", "grid"]
