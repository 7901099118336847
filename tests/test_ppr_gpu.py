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


def _l1(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).sum())


@pytest.mark.parametrize("pull", [1, 50, 1000])
def test_pull_rounds_match_push(gpu, pull):
    # rounds whose frontier holds more than pull/1000 of the out-edges gather
    # over in-edges; the result is the push's up to FP32 summation order
    G = rmat_graph(16, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(5)
    srcs = [int(np.argmax(deg))] + [int(x) for x in rng.choice(np.flatnonzero(deg > 0), 2)]
    with _g(G) as g:
        # the in-edge rows are built at the first host check after a round
        # qualified for pull, so the first query may run all its rounds as pushes
        g.set_tune(ppr_pull=pull)
        g.ppr(srcs[0], alpha=ALPHA, eps=1e-9)
        for s in srcs:
            g.set_tune(ppr_pull=0)
            p0, r0, st0 = g.ppr(s, alpha=ALPHA, eps=1e-9)
            g.set_tune(ppr_pull=pull)
            p1, r1, st1 = g.ppr(s, alpha=ALPHA, eps=1e-9)
            assert st0.pull_rounds == 0
            assert (st1.pull_rounds > 0) == (pull < 1000), (s, st1.pull_rounds)
            assert _l1(p1, p0) <= 2e-5 and _l1(r1, r0) <= 2e-5, s
            pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-15)
            assert abs(_l1(p1, pi) - r1.astype(np.float64).sum()) <= 1e-5
            assert np.all(r1 <= 1e-9 * deg * (1 + 1e-5) + 1e-30)


def test_pull_rounds_directed_and_small(gpu):
    # dangling vertices, vertices without in-edges (no pull items), a graph
    # whose V is not a multiple of 32, and the accuracy bound at small eps
    G = rmat_graph(14, 5, wmax=0, want_f=True)
    s = int(np.argmax(np.diff(G["vlist"])))
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-15)
    with _g(G) as g:
        g.set_tune(ppr_pull=1)
        g.ppr(s, alpha=ALPHA, eps=1e-11)  # builds the in-edge rows
        p, r, st = g.ppr(s, alpha=ALPHA, eps=1e-11)
    assert st.pull_rounds > 0
    assert _l1(p, pi) <= 1e-5
    H = random_graph(301, 2500, seed=6)
    with _g(H) as g:
        g.set_tune(ppr_pull=1)
        g.ppr(0, alpha=ALPHA, eps=1e-9)  # builds the in-edge rows
        g.set_tune(ppr_pull=0)
        for s in (0, 150, 300):
            p0, r0, _ = g.ppr(s, alpha=ALPHA, eps=1e-9)
            g.set_tune(ppr_pull=1)
            p1, r1, st = g.ppr(s, alpha=ALPHA, eps=1e-9)
            g.set_tune(ppr_pull=0)
            assert st.pull_rounds > 0 and _l1(p1, p0) <= 2e-5 and _l1(r1, r0) <= 2e-5, s


def test_pull_rounds_pruned_graph_and_reference_modes(gpu):
    # pruned ppr/ variant (one-hop seeding, then the live edges), the
    # reference's adsorption push, and the budgeted GPU -> host hand-over
    G = random_graph(300, 900, seed=4)
    P = c2g.prune_ppr_mem(G["vlist"], G["elist"], G["flist"])
    with c2g.Graph.from_csr(P["vlist"], P["elist"], flist=P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwf=P["hwlist"], variant=c2g.VARIANT_PPR_PURN, device=0) as g:
        g.set_tune(ppr_pull=1)
        g.ppr(0, alpha=ALPHA, eps=1e-9)  # builds the in-edge rows
        g.set_tune(ppr_pull=0)
        for s in range(0, 300, 37):
            p0, r0, _ = g.ppr(s, alpha=ALPHA, eps=1e-9)
            g.set_tune(ppr_pull=1)
            p1, r1, _ = g.ppr(s, alpha=ALPHA, eps=1e-9)
            g.set_tune(ppr_pull=0)
            assert _l1(p1, p0) <= 2e-5 and _l1(r1, r0) <= 2e-5, s
    H = rmat_graph(14, 3, wmax=0, symmetric=True, want_f=True)
    s = int(np.argmax(np.diff(H["vlist"])))
    with _g(H) as g:
        a0, b0, _ = g.adsorption_push(s)
        m0, n0, sm0 = g.ppr_migrate(s, 3, 1 << 30, alpha=ALPHA, eps=1e-9)
        g.set_tune(ppr_pull=1)
        g.adsorption_push(s)  # builds the in-edge rows
        a1, b1, sa = g.adsorption_push(s)
        m1, n1, sm1 = g.ppr_migrate(s, 3, 1 << 30, alpha=ALPHA, eps=1e-9)
    assert sa.pull_rounds > 0
    # the adsorption push seeds residual 10: ten times the unit-mass bound 2e-5
    assert _l1(a1, a0) <= 2e-4 and _l1(b1, b0) <= 2e-4
    assert _l1(m1, m0) <= 2e-5 and _l1(n1, n0) <= 2e-5


@pytest.mark.slow
def test_c3_full_scale_l1_within_1e5(gpu):
    # the C3 graph and source again, at eps = 1e-13*outdeg: the push's own
    # truncation sum(r) falls below 1e-5, and the FP32 result meets L1 <= 1e-5
    # and max-abs <= 1e-6 against the FP64 power iteration.  Most rounds are
    # dense and run as pull sweeps at the default switch point.
    G = rmat_graph(22, 3, wmax=0, symmetric=True, want_f=True)
    deg = np.diff(G["vlist"])
    rng = np.random.default_rng(3)
    s = int(rng.choice(np.flatnonzero(deg > 0)))
    with _g(G) as g:
        p, r, st = g.ppr(s, alpha=ALPHA, eps=1e-13)
    pi = oracle.ppr_power(G["vlist"], G["elist"], G["flist"], s, alpha=ALPHA, tol=1e-13)
    d = np.abs(p.astype(np.float64) - pi)
    l1, mx, sr = d.sum(), d.max(), r.astype(np.float64).sum()
    print(f"\nC3 scale 22 source {s} eps 1e-13: L1 {l1:.6g} max-abs {mx:.6g} sum(r) {sr:.6g} "
          f"rounds {st.iterations} (pull {st.pull_rounds}) {st.ms:.1f} ms")
    assert st.pull_rounds > 0
    assert l1 <= 1e-5 and mx <= 1e-6
    assert abs(l1 - sr) <= 1e-6
