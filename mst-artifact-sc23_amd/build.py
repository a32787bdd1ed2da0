"""Builds libemst.so (gfx950 HIP kernels + C++ host code) and the CLIs.

Outputs, in-tree so they travel with the source tree:
  mst-artifact-sc23_amd/lib/libemst.so
  mst-artifact-sc23_amd/bin/ecl_mst      reference CLI protocol (ECL_MST.cu:398-455)
  mst-artifact-sc23_amd/bin/ecl_gen      seeded synthetic inputs as .egr files
  mst-artifact-sc23_amd/bin/ecl_convert  .egr to the external baselines' formats
Object files go to mst-artifact-sc23_amd/_build/.
Usage: python mst-artifact-sc23_amd/build.py [-j N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function"]

HOST_SOURCES = ["egr_io.cpp", "generators.cpp", "cpu_mst.cpp", "formats.cpp"]
HIP_SOURCES = ["emst_solver.hip", "learned.hip", "batch.hip", "ladder.hip", "contract.hip", "edgelist.hip",
               "noalloc.hip", "relabel.hip"]
CLI_SOURCES = ["ecl_mst_cli.cpp", "ecl_gen_cli.cpp", "ecl_convert_cli.cpp", "contract_mst_cli.cpp", "edgelist_mst_cli.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = [CXX] + CXX_FLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(jobs=8):
    for d in (OBJ, LIB, BIN):
        os.makedirs(d, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HOST_SOURCES + HIP_SOURCES))
        cli_objs = list(ex.map(_compile, CLI_SOURCES))
    lib = os.path.join(LIB, "libemst.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"])
    os.replace(tmp, lib)
    for name, obj in zip(("ecl_mst", "ecl_gen", "ecl_convert", "contract_mst", "edgelist_mst"), cli_objs):
        _run([CXX, "-o", os.path.join(BIN, name), obj, f"-L{LIB}", "-lemst", "-Wl,-rpath,$ORIGIN/../lib", "-pthread"])
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(a.j))
