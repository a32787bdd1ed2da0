"""Records SHA-256 digests of small generated graphs (tests/golden/gen_hashes.json).

The workload generators define the benchmark inputs (inMST positions j are
CSR positions), so their bytes are pinned: a change to a generator shows up
as a digest mismatch in tests/test_host.py.  Run after an intentional
generator change: python tests/golden/make_gen_hashes.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "mst-artifact-sc23_amd"))
import emst  # noqa: E402

CASES = {
    "grid_16x24_s1": lambda: emst.gen_grid(16, 24, 1, 1, 1 << 20),
    "rmat_10_16_s2": lambda: emst.gen_rmat(10, 16, 0.57, 0.19, 0.19, 2, 1, 1 << 24),
    "rmat_9_16_s5_ties": lambda: emst.gen_rmat(9, 16, 0.57, 0.19, 0.19, 5, 1, 65),
    "rgg_4096_s3": lambda: emst.gen_rgg(4096, 6.0, 3, 1 << 24),
    "er_3000_s4": lambda: emst.gen_er(3000, 10.0, 4, 1, 1 << 20),
}


def digest(g):
    h = hashlib.sha256()
    for a in (g.nindex, g.nlist, g.eweight):
        h.update(a.tobytes())
    return h.hexdigest()


if __name__ == "__main__":
    out = {k: digest(f()) for k, f in CASES.items()}
    with open(os.path.join(HERE, "gen_hashes.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)
