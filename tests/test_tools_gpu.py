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
    for col in ("ecl_mst", "ecl_mst_learned", "UMinho_gpu", "Jucele_gpu"):
        cap = (tmp_path / f"{col}_out.csv").read_text()
        assert "ERROR" not in cap and cap.count("input:") == 2
        for p in (str(egr), str(grid)):
            assert runtime[col][p] > 0 and thru[col][p] > 0
    lines = (tmp_path / "comparison_throughputs.csv").read_text().splitlines()
    assert lines[0] == ",ecl_mst, ecl_mst_learned, UMinho_gpu, Jucele_gpu, "
    assert [ln.split(",")[0] for ln in lines[1:]] == ["rmat12", "grid"]


def test_cli_deopt_ladder(tmp_path):
    """Every rung through the CLI (EMST_VARIANT) with the reference's
    verification, and the deopt tables of generate_deoptimize_tables.py."""
    import tables
    egr = tmp_path / "rmat11.egr"
    r = subprocess.run([os.path.join(PKG, "bin", "ecl_gen"), "rmat", "11", "8", "0.57", "0.19", "0.19", "3", "1",
                        "1000", str(egr)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("step: tables.deopt", flush=True)
    runtime, _ = tables.deopt(str(tmp_path), [str(egr)])
    cols = ["ecl_mst" + (f"_{v}" if v else "") for v in tables.DEOPT_VARIANTS]
    for col in cols:
        cap = (tmp_path / f"{col}_out.csv").read_text()
        assert "ERROR" not in cap and cap.count("input:") == 1, col
        assert runtime[col][str(egr)] > 0
    lines = (tmp_path / "deopt_throughputs.csv").read_text().splitlines()
    assert lines[0] == "," + "".join(f"{c}, " for c in cols)
    bad = subprocess.run([CLI_PATH, str(egr)], capture_output=True, text=True,
                         env={**os.environ, "EMST_VARIANT": "NAG_XYZ"})
    assert bad.returncode != 0 and "bad EMST_VARIANT" in bad.stderr


CLI_PATH = os.path.join(PKG, "bin", "ecl_mst")
