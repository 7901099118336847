"""Writes tests/golden/prune_*.npz: inputs and the outputs of the reference's
own purn-sssp / purn-ppr (src/purn-sssp.cpp, src/purn-ppr.cpp), compiled from
the reference sources by oracle/Makefile into oracle/_ref/.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_prune_golden.py
The reference's sssp/hvlist.bin is in prune order (one entry per processed
vertex, src/purn-sssp.cpp:283-284) and is stored as written.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "c-2graph_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import c2g  # noqa: E402
from conftest import random_graph  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def appendix_graph():
    # a=0, s=1, v=2, x=3: a->s:1, s->v:1, v->x:1, a->v:5 (one-level seeding misses x from s)
    vl = np.array([0, 2, 3, 4, 4], np.int32)
    el = np.array([1, 2, 2, 3], np.int32)
    wl = np.array([1, 5, 1, 1], np.int32)
    fl = np.array([0.5, 0.5, 1.0, 1.0], np.float32)
    return dict(vlist=vl, elist=el, wlist=wl, flist=fl)


def cases():
    G = c2g.rmat(10, 16, seed=2, wmax=255, want_f=True)
    yield "rmat10", G
    for seed in (1, 2, 3):
        yield f"sparse{seed}", random_graph(200, 420, seed=seed, wmax=9)
    yield "appendix", appendix_graph()


def main():
    out = os.path.join(ROOT, "tests", "golden")
    for name, G in cases():
        with tempfile.TemporaryDirectory() as ds:
            c2g.write_dataset(ds, G["vlist"], G["elist"], G["wlist"], G["flist"])
            for tool in ("purn-sssp", "purn-ppr"):
                subprocess.run([os.path.join(REF, tool), ds], check=True, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL)
            arrs = {f"in_{k}": np.asarray(G[k]) for k in ("vlist", "elist", "wlist", "flist")}
            for sub in ("sssp", "ppr"):
                for f in ("vlist", "elist", "wlist", "hvlist", "helist", "hwlist"):
                    dt = np.float32 if sub == "ppr" and f in ("wlist", "hwlist") else np.int32
                    arrs[f"{sub}_{f}"] = np.fromfile(os.path.join(ds, sub, f + ".bin"), dtype=dt)
        np.savez_compressed(os.path.join(out, f"prune_{name}.npz"), **arrs)
        print(name, {k: v.shape for k, v in arrs.items() if k.startswith("sssp_h")})


if __name__ == "__main__":
    main()
