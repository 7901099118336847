import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("c-2graph_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

INT_MAX = np.iinfo(np.int32).max


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through libc2g.so on cuda:0)")
    config.addinivalue_line("markers", "slow: large-graph case")


def _gpu_ok():
    import c2g
    try:
        return c2g.device_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips only when no device exists; a present device whose library fails is an error."""
    if not _gpu_ok():
        pytest.skip("no gfx950 device")
    return True


_graphs = {}


def rmat_graph(scale, seed, wmax=64, ef=16, symmetric=False, want_f=False):
    import c2g
    key = (scale, seed, wmax, ef, symmetric, want_f)
    if key not in _graphs:
        _graphs[key] = c2g.rmat(scale, ef, seed=seed, wmax=wmax, symmetric=symmetric, want_f=want_f)
    return _graphs[key]


def random_graph(V, E, seed, wmax=10, zero_ok=False):
    """Small random digraph in the reference CSR layout (rows sorted, no loops, no dups)."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, V, E)
    v = rng.integers(0, V, E)
    keep = u != v
    key = np.unique(u[keep].astype(np.int64) * V + v[keep])
    u, v = key // V, key % V
    vl = np.zeros(V + 1, np.int32)
    np.add.at(vl, u + 1, 1)
    vl = np.cumsum(vl).astype(np.int32)
    w = rng.integers(0 if zero_ok else 1, wmax + 1, len(v)).astype(np.int32)
    f = rng.random(len(v)).astype(np.float64) + 1e-3
    rs = np.zeros(V)
    np.add.at(rs, u, f)
    f = (f / rs[u]).astype(np.float32)
    return dict(V=V, E=len(v), vlist=vl, elist=v.astype(np.int32), wlist=w, flist=f)


def no_dangling_graph(V, E, seed):
    """Random digraph with the cycle i -> i+1 added, so no vertex is dangling;
    row-normalised float weights."""
    rng = np.random.default_rng(seed)
    u = np.concatenate([np.arange(V), rng.integers(0, V, E)])
    v = np.concatenate([(np.arange(V) + 1) % V, rng.integers(0, V, E)])
    keep = u != v
    key = np.unique(u[keep].astype(np.int64) * V + v[keep])
    u, v = key // V, key % V
    vl = np.concatenate([[0], np.cumsum(np.bincount(u, minlength=V))]).astype(np.int32)
    f = rng.random(len(v)) + 1e-3
    f = (f / np.bincount(u, weights=f, minlength=V)[u]).astype(np.float32)
    return dict(V=V, E=len(v), vlist=vl, elist=v.astype(np.int32), flist=f)


def site_graph(scale, seed):
    """The web-like site graph of bench.py (where purn-sssp / purn-ppr drop
    about a quarter of the edges)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    return bench.site_graph(scale, seed=seed)


def check_sssp_certificate(vl, el, wl, src, d):
    """Exactness certificate for int SSSP with positive weights (Graph500-style):
    d[s]=0; d[v] <= d[u]+w on every edge out of a reached u; every reached v != s
    has a tight in-edge; unreached vertices are INT_MAX."""
    V = len(vl) - 1
    d64 = d.astype(np.int64)
    assert d[src] == 0
    deg = np.diff(vl)
    u = np.repeat(np.arange(V, dtype=np.int64), deg)
    fin_u = d64[u] != INT_MAX
    cand = d64[u] + wl
    assert np.all(d64[el][fin_u] <= cand[fin_u]), "feasibility violated"
    tight = np.zeros(V, bool)
    t = fin_u & (d64[el] == cand)
    tight[el[t]] = True
    reached = d64 != INT_MAX
    reached[src] = False
    assert np.all(tight[reached]), "a reached vertex has no tight in-edge"


def headline_sources(vlist, n, seed):
    """The batched workload's source recipe: n/2 drawn from the top-1%
    out-degree vertices, n/2 from out-degree-1 vertices (seeded)."""
    deg = np.diff(vlist)
    rng = np.random.default_rng(seed)
    order = np.argsort(-deg, kind="stable")
    top = order[: max(1, len(deg) // 100)]
    ones = np.flatnonzero(deg == 1)
    a = rng.choice(top, n // 2, replace=False)
    b = rng.choice(ones, n - n // 2, replace=len(ones) < n - n // 2)
    return np.concatenate([a, b]).astype(np.int32)


def checksum(d):
    fin = d != INT_MAX
    return int(d[fin].astype(np.int64).sum()), int(fin.sum())
