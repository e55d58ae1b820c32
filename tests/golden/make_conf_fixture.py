"""Extracts a compact known-answer fixture from the reference's example data.

Reads pyhea's examples/conf.lmp (a converged 4000-site FCC 10^3 structure,
atom types 1..4 with 1600/1600/400/400 atoms, Cartesian positions in units of
a/2 = 1.8076171875 A) and examples/TARGET_SRO (first-shell target, lower
triangle), and writes tests/golden/conf_fcc10.npz:
  positions  int16 [4000][3]  half-lattice units, file order
  types      uint8 [4000]     atom type - 1, file order
  target1    f64   [4][4]     first-shell target alpha
Only data is stored.

    python tests/golden/make_conf_fixture.py [/path/to/pyhea]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def main():
    pos, typ = [], []
    with open(os.path.join(REF, "examples", "conf.lmp")) as f:
        for line in f:
            p = line.split()
            if len(p) == 5 and p[0].isdigit():
                typ.append(int(p[1]) - 1)
                pos.append([float(v) for v in p[2:]])
    half = np.array(pos) / 1.8076171875
    ih = np.rint(half).astype(np.int16)
    assert np.abs(half - ih).max() < 1e-6
    tri = []
    with open(os.path.join(REF, "examples", "TARGET_SRO")) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            try:
                tri.extend(float(v) for v in s.split())
            except ValueError:
                continue
    K = 4
    t = np.zeros((K, K))
    for i in range(K):
        for j in range(K):
            hi, lo = max(i, j), min(i, j)
            t[i, j] = tri[hi * (hi + 1) // 2 + lo]
    np.savez_compressed(os.path.join(HERE, "conf_fcc10.npz"), positions=ih,
                        types=np.array(typ, np.uint8), target1=t)
    print("wrote conf_fcc10.npz", ih.shape, np.bincount(typ))


if __name__ == "__main__":
    main()
