"""Writes tests/golden/stanford_layout.npz from the only data the reference
ships with its sources: dataset/stanford/{origin,sssp,ppr}/{vlist,hvlist}.bin
(a web-BerkStan-sized graph; the edge files are not shipped).  Each array is
stored as its first differences (row degrees), which compress well.

Run from the repository root:
    python tests/golden/make_stanford_fixture.py
"""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SRC = "/root/reference/dataset/stanford"
FILES = ["origin/vlist.bin", "sssp/vlist.bin", "sssp/hvlist.bin", "ppr/vlist.bin", "ppr/hvlist.bin"]


def main():
    out = {}
    for f in FILES:
        a = np.fromfile(os.path.join(SRC, f), dtype=np.int32)
        assert a[0] == 0 and np.all(np.diff(a) >= 0), f
        out[f.replace("/", "_").replace(".bin", "")] = np.diff(a).astype(np.int32)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "stanford_layout.npz"), **out)


if __name__ == "__main__":
    main()
