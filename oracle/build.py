"""Builds the oracle (test infrastructure only) and, where the reference
tree is present, the reference-side checkers in oracle/_ref/.

  oracle/_build/liboracle.so   gcc build of emst_oracle.c + cpu_baselines.c
  oracle/_ref/ref_egr_dump     g++ build of oracle/ref_egr_dump.cpp against the
                               reference's own ECLgraph.h (read in place from
                               /root/reference/ECL-MST, never copied)
  oracle/_ref/ecl2pbbs, ecl2lsg, ecl2mtx, ecl2uminho
                               the reference's format converters
                               (*/inputs/ecl2*_genweights.cpp), compiled in
                               place with each directory's own ECLgraph.h, as
                               the reference's build scripts do
                               (PBBS_CPU/build_PBBS.sh:15 and the others, minus
                               -march=native)

The rest of the reference's hot path (ECL_MST.cu) needs nvcc and <cuda.h>,
neither of which exists here, so it is not built.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ECL = "/root/reference/ECL-MST"
REF_ROOT = "/root/reference"
# converter name -> (directory under the reference tree, source file)
REF_CONVERTERS = {
    "ecl2pbbs": ("PBBS_CPU/inputs", "ecl2pbbs_genweights.cpp"),
    "ecl2lsg": ("Lonestar_CPU/inputs", "ecl2lsg_genweights.cpp"),
    "ecl2mtx": ("cuGraph_GPU/inputs", "ecl2mtx_genweights_cugraph.cpp"),
    "ecl2uminho": ("UMinho_BOTH/inputs", "ecl2UMinho_genweights.cpp"),
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError("oracle build failed")


def build():
    out = os.path.join(HERE, "_build")
    os.makedirs(out, exist_ok=True)
    so = os.path.join(out, "liboracle.so")
    srcs = [os.path.join(HERE, f) for f in ("emst_oracle.c", "cpu_baselines.c")
            if os.path.exists(os.path.join(HERE, f))]
    _run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-shared", "-fopenmp", "-o", so + ".tmp"]
         + srcs + ["-lm"])
    os.replace(so + ".tmp", so)
    built = [so]
    if os.path.exists(os.path.join(REF_ECL, "ECLgraph.h")):
        refdir = os.path.join(HERE, "_ref")
        os.makedirs(refdir, exist_ok=True)
        exe = os.path.join(refdir, "ref_egr_dump")
        _run(["g++", "-O2", "-I", REF_ECL, "-o", exe, os.path.join(HERE, "ref_egr_dump.cpp")])
        built.append(exe)
        for name, (d, src) in REF_CONVERTERS.items():
            path = os.path.join(REF_ROOT, d, src)
            if os.path.exists(path):
                exe = os.path.join(refdir, name)
                _run(["g++", "-O3", "-w", "-I", os.path.join(REF_ROOT, d), "-o", exe, path])
                built.append(exe)
    return built


if __name__ == "__main__":
    print("\n".join(build()))
