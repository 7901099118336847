"""The seeded R-MAT generator writes the reference's CSR layout
(utils/convert2binary.cpp:67,76-79: no self-loops, rows sorted, no duplicates)."""
import numpy as np

import c2g


def test_layout_and_determinism():
    G = c2g.rmat(14, 16, seed=3, wmax=255, want_f=True)
    H = c2g.rmat(14, 16, seed=3, wmax=255, want_f=True)
    for k in ("vlist", "elist", "wlist", "flist"):
        assert np.array_equal(G[k], H[k])
    V, vl, el = G["V"], G["vlist"], G["elist"]
    assert vl[0] == 0 and vl[-1] == G["E"] == len(el) and np.all(np.diff(vl) >= 0)
    u = np.repeat(np.arange(V), np.diff(vl))
    assert np.all(u != el)
    key = u.astype(np.int64) * V + el
    assert np.all(np.diff(key) > 0)
    assert G["wlist"].min() >= 1 and G["wlist"].max() <= 255
    rs = np.zeros(V)
    np.add.at(rs, u, G["flist"].astype(np.float64))
    has = np.diff(vl) > 0
    assert np.abs(rs[has] - 1).max() < 1e-5 and np.all(G["flist"] > 0)


def test_seed_changes_graph_and_weights_do_not_depend_on_wmax_layout():
    a = c2g.rmat(12, 8, seed=1, wmax=64)
    b = c2g.rmat(12, 8, seed=2, wmax=64)
    assert a["E"] != b["E"] or not np.array_equal(a["elist"], b["elist"])
    c = c2g.rmat(12, 8, seed=1, wmax=0)
    assert c["wlist"] is None and np.array_equal(c["elist"], a["elist"])


def test_symmetric():
    G = c2g.rmat(12, 8, seed=5, wmax=10, symmetric=True)
    V = G["V"]
    u = np.repeat(np.arange(V), np.diff(G["vlist"]))
    fwd = set(zip(u.tolist(), G["elist"].tolist()))
    assert all((v, x) in fwd for x, v in fwd)


def test_skew_and_permutation():
    # quadrant a=0.57 gives a heavy-tailed out-degree; the vertex bijection
    # spreads hubs over the id range instead of piling them at id 0
    G = c2g.rmat(16, 16, seed=1, wmax=1)
    deg = np.diff(G["vlist"])
    assert deg.max() > 50 * deg.mean()
    top = np.argsort(-deg)[:20]
    assert top.max() > G["V"] // 4
    # same multiset of degrees for any seed's bijection is not required, but
    # every vertex id is in range and the edge count is near ef*V minus dups
    assert 0.5 * 16 * G["V"] < G["E"] <= 16 * G["V"]
