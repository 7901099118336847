"""Forward-push PPR on the GPU (calcuatePPR<float>, src/ppr/base/Main.cu:121-388;
pruned seeding src/ppr/purn/Main.cu:78-101) against the oracle.

Tolerances (FP32 state, atomics in arbitrary order):
  * certificate: for a graph without dangling vertices, the exact PPR vector
    satisfies |p - pi|_1 = sum(r) for any push state; we require
    | |p - pi|_1 - sum(r) | <= 1e-5 and r_v <= eps*deg(v) at termination;
  * accuracy: with eps small enough that the push's own error is below it,
    |p - pi|_1 <= 1e-5 and max|p - pi| <= 1e-6 (pi = FP64 power iteration);
  * same algorithm: |p_gpu - p_oracle_push|_1 <= 2e-5 at eps = 1e-7*deg.
"""
import numpy as np
import pytest

import c2g
import oracle
from conftest import random_graph, rmat_graph

pytestmark = pytest.mark.gpu
ALPHA = 0.15


def _g(G):
    return c2g.Graph.from_csr(G["vlist"], G["elist"], flist=G["flist"], device=0)


def test_certificate_and_termination_bound(gpu):
    G = rmat_graph(16, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    s = int(np.argmax(deg))
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-15)
    with _g(G) as g:
        p, r, st = g.ppr(s, alpha=ALPHA, eps=1e-7)
    assert st.lane == c2g.LANE_GPU and st.iterations > 0
    assert np.all(r <= 1e-7 * deg * (1 + 1e-5) + 1e-30)
    l1 = np.abs(p.astype(np.float64) - pi).sum()
    assert abs(l1 - r.astype(np.float64).sum()) <= 1e-5, (l1, r.sum())
    assert abs(p.sum() + r.sum() - 1.0) <= 1e-4


def test_accuracy_at_small_threshold(gpu):
    G = rmat_graph(14, 3, wmax=0, symmetric=True, want_f=True)
    s = int(np.argmax(np.diff(G["vlist"])))
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-15)
    with _g(G) as g:
        p, r, _ = g.ppr(s, alpha=ALPHA, eps=1e-11)
    d = np.abs(p.astype(np.float64) - pi)
    assert d.sum() <= 1e-5, d.sum()
    assert d.max() <= 1e-6, d.max()


def test_matches_serial_push(gpu):
    G = rmat_graph(14, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(1)
    with _g(G) as g:
        for s in [int(np.argmax(deg))] + [int(x) for x in rng.choice(np.flatnonzero(deg > 0), 3)]:
            po, ro, _, _ = oracle.ppr_push(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, eps=1e-7)
            p, r, _ = g.ppr(s, alpha=ALPHA, eps=1e-7)
            assert np.abs(p.astype(np.float64) - po).sum() <= 2e-5, s
            assert np.abs(r.astype(np.float64) - ro).sum() <= 2e-5, s


def test_dangling_and_directed(gpu):
    # directed graph with dangling vertices: mass is dropped like the reference
    G = rmat_graph(14, 5, wmax=0, want_f=True)
    s = int(np.argmax(np.diff(G["vlist"])))
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-15)
    with _g(G) as g:
        p, r, _ = g.ppr(s, alpha=ALPHA, eps=1e-11)
    assert np.abs(p.astype(np.float64) - pi).sum() <= 1e-5
    with _g(G) as g:
        with pytest.raises(c2g.C2GError):
            g.ppr(s, alpha=1.5)


def test_batched_equals_single(gpu):
    G = rmat_graph(14, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(2)
    srcs = rng.choice(np.flatnonzero(deg > 0), 70).astype(np.int32)  # 2 groups
    with _g(G) as g:
        P, R, st = g.ppr_batch(srcs, alpha=ALPHA, eps=1e-7)
        for q in (0, 1, 63, 64, 69):
            p, r, _ = g.ppr(int(srcs[q]), alpha=ALPHA, eps=1e-7)
            assert np.abs(P[q].astype(np.float64) - p).sum() <= 2e-5, q
    for q in (0, 69):
        pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], int(srcs[q]), alpha=ALPHA, tol=1e-15)
        l1 = np.abs(P[q].astype(np.float64) - pi).sum()
        assert abs(l1 - R[q].astype(np.float64).sum()) <= 1e-5


def test_batched_ragged_sizes(gpu):
    # V not a multiple of the 64-vertex extraction tile; 14 queries = one full
    # and one partial 8-column transfer; r columns come out too
    G = random_graph(301, 2500, seed=6)
    srcs = np.arange(0, 301, 23, dtype=np.int32)
    with _g(G) as g:
        P, R, _ = g.ppr_batch(srcs, alpha=ALPHA, eps=1e-7)
        for q, s in enumerate(srcs):
            p, r, _ = g.ppr(int(s), alpha=ALPHA, eps=1e-7)
            assert np.abs(P[q].astype(np.float64) - p).sum() <= 2e-5, q
            assert np.abs(R[q].astype(np.float64) - r).sum() <= 2e-5, q


def test_pruned_graph_seeding(gpu):
    # ppr/ variant: the first frontier is H(s) U N'(s) with r = (1-alpha) w; the
    # oracle runs the same push from that seeded state on the pruned graph.
    G = random_graph(300, 900, seed=4)
    P = c2g.prune_ppr_mem(G["vlist"], G["elist"], G["flist"])
    assert len(P["helist"]) > 0
    with c2g.Graph.from_csr(P["vlist"], P["elist"], flist=P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwf=P["hwlist"], variant=c2g.VARIANT_PPR_PURN, device=0) as g:
        for s in range(0, 300, 37):
            hv, hl, hw = P["hvlist"], P["helist"], P["hwlist"]
            vl, el, wl = P["vlist"], P["elist"], P["wlist"]
            nb = np.concatenate([hl[hv[s]:hv[s + 1]], el[vl[s]:vl[s + 1]]]).astype(np.int32)
            ww = np.concatenate([hw[hv[s]:hv[s + 1]], wl[vl[s]:vl[s + 1]]]).astype(np.float32)
            p0 = np.zeros(300, np.float32)
            r0 = np.zeros(300, np.float32)
            p0[s] = ALPHA
            np.add.at(r0, nb, (np.float32(1 - ALPHA) * ww).astype(np.float32))
            fr = np.unique(nb).astype(np.int32)
            p0[fr] += r0[fr] * np.float32(ALPHA)
            if len(fr):
                oracle.lib().orc_forward_push(300, vl, el, wl, ALPHA, 1e-7, p0, r0, fr, len(fr), 1 << 30,
                                              np.zeros(1, np.int64))
            p, r, _ = g.ppr(s, alpha=ALPHA, eps=1e-7)
            assert np.abs(p - p0).sum() <= 2e-5, s


def test_pruned_graph_batched_equals_single(gpu):
    # batched seeding on the ppr/ variant (H(s) U N'(s) with r = (1-alpha) w)
    # must give the single-source results for every lane
    G = random_graph(300, 900, seed=7)
    P = c2g.prune_ppr_mem(G["vlist"], G["elist"], G["flist"])
    assert len(P["helist"]) > 0
    srcs = np.arange(0, 300, 7, dtype=np.int32)
    with c2g.Graph.from_csr(P["vlist"], P["elist"], flist=P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwf=P["hwlist"], variant=c2g.VARIANT_PPR_PURN, device=0) as g:
        Pb, Rb, _ = g.ppr_batch(srcs, alpha=ALPHA, eps=1e-7)
        for q, s in enumerate(srcs):
            p, r, _ = g.ppr(int(s), alpha=ALPHA, eps=1e-7)
            assert np.abs(Pb[q].astype(np.float64) - p).sum() <= 2e-5, s


@pytest.mark.slow
def test_c3_full_scale_against_power_iteration(gpu):
    # the PPR workload at full size: symmetrised R-MAT scale 22, seed 3,
    # alpha 0.15, eps 1e-7*outdeg, one source, against the FP64 power
    # iteration.  The certificate |p - pi|_1 = sum(r) holds for the push
    # whatever eps is; L1 and max-abs are printed for the report.
    G = rmat_graph(22, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    # symmetrised: a vertex without out-edges has no in-edges either, so no
    # reachable vertex is dangling and no mass is dropped
    assert np.all(np.bincount(G["elist"], minlength=G["V"])[deg == 0] == 0)
    rng = np.random.default_rng(3)
    s = int(rng.choice(np.flatnonzero(deg > 0)))
    with _g(G) as g:
        p, r, st = g.ppr(s, alpha=ALPHA, eps=1e-7)
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-13)
    d = np.abs(p.astype(np.float64) - pi)
    l1, mx, sr = d.sum(), d.max(), r.astype(np.float64).sum()
    print(f"\nC3 scale 22 source {s}: L1 {l1:.6g} max-abs {mx:.6g} sum(r) {sr:.6g} rounds {st.iterations}")
    assert abs(l1 - sr) <= 1e-5
    assert np.all(r <= deg * 1e-7 + 1e-12)
