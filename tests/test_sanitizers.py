"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):
the library's host sources (reader, writers, generators, validation,
threshold, Kruskal, verify) and the oracle are built with
-fsanitize=address,undefined and run by tests/sanitize_driver.cpp on the
reader fixtures, generated graphs, every writer and 300 byte-mutated files.
Any sanitizer report or any disagreement with the oracle fails."""
import os
import shutil
import subprocess

import pytest

from test_oracle import egr_fixtures

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "mst-artifact-sc23_amd", "csrc")
HOST = ["egr_io.cpp", "generators.cpp", "cpu_mst.cpp", "formats.cpp"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(not shutil.which("g++") or not shutil.which("gcc"), reason="no host compiler")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    objs = []
    for f in ("emst_oracle.c", "cpu_baselines.c"):
        o = str(tmp_path / (f + ".o"))
        subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-fopenmp", "-c", os.path.join(ROOT, "oracle", f), "-o", o]
                       + SAN, check=True, capture_output=True)
        objs.append(o)
    exe = str(tmp_path / "sanitize_driver")
    srcs = [os.path.join(CSRC, f) for f in HOST] + [os.path.join(HERE, "sanitize_driver.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-pthread", "-fopenmp", "-o", exe] + SAN + srcs + objs + ["-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    fx = tmp_path / "fixtures"
    fx.mkdir()
    egr_fixtures(str(fx))
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=23:verify_asan_link_order=0",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1", "OMP_NUM_THREADS": "2"}
    r = subprocess.run([exe, str(fx), str(scratch)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "failures 0" in r.stdout
