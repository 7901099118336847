"""Every workload at its own size on the GPU, against the oracle.

  C4  batched SSSP: R-MAT scale 24 seed 4, w U[1,1023], 256 sources (128 from
      the top-1% out-degree vertices, 128 of out-degree 1)
  C2  pruned SSSP: R-MAT scale 22 seed 2, w U[1,255], purn-sssp, 32 uniform
      sources, single-source and batched, against full-graph Dijkstra
  C3  PPR: symmetrised R-MAT scale 22 seed 3, 64 sources batched
  C5  served mixed stream on R-MAT scale 23 seed 5 at 40 QPS
  C1  sssp-purn on the seed-1 scale-16 graph (w U[1,64])

SSSP semantics follow CalcuSSSP (src/sssp/base/Main.cu:109-356) and the pruned
seeding of src/sssp/purn/Main.cu:67-90; exactness is checked against Dijkstra
and, where Dijkstra would be too slow for every result, by the oracle's
certificate, which proves a distance vector exact for weights >= 1.
"""
import time

import numpy as np
import pytest

import c2g
import oracle
from conftest import INT_MAX, checksum, headline_sources, rmat_graph

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
ALPHA = 0.15


def say(*a):
    print("[fullsize]", *a, flush=True)


def test_c4_batched_scale24_256_sources(gpu):
    t = time.time()
    G = rmat_graph(24, 4, wmax=1023)
    vl, el, wl = G["vlist"], G["elist"], G["wlist"]
    srcs = headline_sources(vl, 256, seed=4)
    say(f"C4 graph V={G['V']} E={G['E']} in {time.time() - t:.1f}s")
    with c2g.Graph.from_csr(vl, el, wl, device=0) as g:
        sums_b, reach_b, st_b = g.sssp_batch_resident(srcs)
        say(f"C4 batched: rounds {st_b.iterations} lane relaxations {st_b.edges_relaxed}")
        sums_e, reach_e, _ = g.sssp_each_resident(srcs)
        assert np.array_equal(sums_b, sums_e) and np.array_equal(reach_b, reach_e)
        # 4 hubs and 4 out-degree-1 sources against Dijkstra
        pick = np.concatenate([np.arange(0, 4), np.arange(128, 132)])
        sample = [int(srcs[k]) for k in pick]
        singles = [g.sssp(s)[0] for s in sample]
        d8, _ = g.sssp_batch(np.array(sample, np.int32))  # the batched kernels' full vectors
    refs = oracle.dijkstra_many(vl, el, wl, sample)
    for k, s, d1, db, ref in zip(pick, sample, singles, d8, refs):
        assert checksum(ref) == (int(sums_b[k]), int(reach_b[k])), s
        assert np.array_equal(d1, ref), s
        assert np.array_equal(db, ref), s
        assert oracle.sssp_certificate(vl, el, wl, s, d1) == 0, s
    say("C4 ok")


def test_c2_pruned_scale22_32_sources(gpu):
    G = rmat_graph(22, 2, wmax=255)
    vl, el, wl = G["vlist"], G["elist"], G["wlist"]
    t = time.time()
    P = c2g.prune_sssp_mem(vl, el, wl)
    say(f"C2 pruned in {time.time() - t:.1f}s: kept {len(P['elist'])}/{G['E']}, hidden {len(P['helist'])}")
    deg = np.diff(vl)
    rng = np.random.default_rng(2)
    srcs = rng.choice(np.flatnonzero(deg > 0), 32, replace=False).astype(np.int32)
    refs = oracle.dijkstra_many(vl, el, wl, srcs)
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
        db, _ = g.sssp_batch(srcs)
        for q, s in enumerate(srcs):
            d, _ = g.sssp(int(s))
            assert np.array_equal(d, refs[q]), int(s)
            assert np.array_equal(db[q], refs[q]), (int(s), "batched")
    say("C2 ok")


def test_c1_sssp_purn_seed1(gpu):
    # C1's graph and source (max out-degree), pruned, against full-graph Dijkstra
    G = rmat_graph(16, 1, wmax=64)
    vl, el, wl = G["vlist"], G["elist"], G["wlist"]
    P = c2g.prune_sssp_mem(vl, el, wl)
    s = int(np.argmax(np.diff(vl)))
    srcs = np.concatenate([[s], P["porder"][:15]]).astype(np.int32)
    with c2g.Graph.from_csr(P["vlist"], P["elist"], P["wlist"], hvlist=P["hvlist"], helist=P["helist"],
                            hwlist=P["hwlist"], variant=c2g.VARIANT_SSSP_PURN, device=0) as g:
        db, _ = g.sssp_batch(srcs)
        for q, x in enumerate(srcs):
            ref = oracle.dijkstra(vl, el, wl, int(x))
            d, _ = g.sssp(int(x))
            assert np.array_equal(d, ref) and np.array_equal(db[q], ref), int(x)


def test_c3_batched_scale22_64_sources(gpu):
    G = rmat_graph(22, 3, wmax=0, symmetric=True, want_f=True)
    vl, el, fl = G["vlist"], G["elist"], G["flist"]
    deg = np.diff(vl)
    rng = np.random.default_rng(3)
    srcs = rng.choice(np.flatnonzero(deg > 0), 64, replace=False).astype(np.int32)
    with c2g.Graph.from_csr(vl, el, flist=fl, device=0) as g:
        pb, rb, st = g.ppr_batch(srcs, alpha=ALPHA, eps=1e-7)
        say(f"C3 batched: rounds {st.iterations}")
        worst = 0.0
        for q, s in enumerate(srcs):
            p, r, _ = g.ppr(int(s), alpha=ALPHA, eps=1e-7)
            l1 = float(np.abs(pb[q].astype(np.float64) - p).sum())
            worst = max(worst, l1)
            assert l1 <= 2e-5, (int(s), l1)
            assert np.all(rb[q] <= deg * 1e-7 * (1 + 1e-5) + 1e-30), int(s)
    say(f"C3 worst lane-vs-single L1 {worst:.3g}")
    # the push certificate |p - pi|_1 = sum(r) for two lanes (no reachable
    # vertex is dangling on a symmetrised graph)
    for q in (0, 1):
        pi = oracle.ppr_power(vl, el, fl, int(srcs[q]), alpha=ALPHA, tol=1e-13)
        l1 = float(np.abs(pb[q].astype(np.float64) - pi).sum())
        sr = float(rb[q].astype(np.float64).sum())
        say(f"C3 lane {q} source {int(srcs[q])}: L1 {l1:.6g} sum(r) {sr:.6g}")
        assert abs(l1 - sr) <= 1e-5


def test_c5_served_stream_scale23(gpu):
    # 5 s of Poisson arrivals at 40 QPS, 60% SSSP / 30% PPR / 10% adsorption;
    # every result is checked after the stream: SSSP exact (certificate for
    # all, Dijkstra for 8), PPR and adsorption against the query run alone
    G = rmat_graph(23, 5, wmax=255, want_f=True)
    vl, el, wl, fl = G["vlist"], G["elist"], G["wlist"], G["flist"]
    deg = np.diff(vl)
    pool = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(5)
    qps, secs, L = 40.0, 5.0, 16
    n = int(qps * secs)
    kinds = rng.choice([c2g.Q_SSSP, c2g.Q_PPR, c2g.Q_ADS], n, p=[0.6, 0.3, 0.1])
    srcs = np.where(kinds == c2g.Q_ADS, rng.integers(0, 8, n), rng.choice(pool, n))
    gaps = rng.exponential(1.0 / qps, n)
    with c2g.Graph.from_csr(vl, el, wl, fl, device=0) as g:
        results = []
        with c2g.Server(g, max_batch=64, ads_labels=L, ads_iters=20, alpha=ALPHA, eps=1e-7) as s:
            ids = []
            t0 = time.perf_counter()
            nxt = t0
            for i in range(n):
                nxt += gaps[i]
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append(s.submit(int(kinds[i]), int(srcs[i])))
            say(f"C5 submitted {n} in {time.perf_counter() - t0:.1f}s")
            for i, rid in enumerate(ids):
                info, res = s.poll(rid)
                assert info["status"] == 0 and info["lane"] == c2g.LANE_GPU, info
                results.append(res)
        say("C5 stream done; checking")
        sssp_idx = [i for i in range(n) if kinds[i] == c2g.Q_SSSP]
        for i in sssp_idx:
            assert oracle.sssp_certificate(vl, el, wl, int(srcs[i]), results[i]) == 0, int(srcs[i])
        refs = oracle.dijkstra_many(vl, el, wl, [int(srcs[i]) for i in sssp_idx[:8]])
        for i, ref in zip(sssp_idx[:8], refs):
            assert np.array_equal(results[i], ref), int(srcs[i])
        for i in range(n):
            if kinds[i] == c2g.Q_PPR:
                p, _, _ = g.ppr(int(srcs[i]), alpha=ALPHA, eps=1e-7)
                assert np.abs(results[i].astype(np.float64) - p).sum() <= 2e-5, int(srcs[i])
        # each adsorption seed set run alone, one request on a fresh server
        alone = {}
        with c2g.Server(g, max_batch=64, ads_labels=L, ads_iters=20, alpha=ALPHA, eps=1e-7) as s:
            for k in sorted(set(int(srcs[i]) for i in range(n) if kinds[i] == c2g.Q_ADS)):
                alone[k] = s.poll(s.submit(c2g.Q_ADS, k))[1]
        for i in range(n):
            if kinds[i] == c2g.Q_ADS:
                assert np.abs(results[i] - alone[int(srcs[i])]).max() <= 2e-6, int(srcs[i])
    say(f"C5 ok: {len(sssp_idx)} SSSP, {int((kinds == c2g.Q_PPR).sum())} PPR, {int((kinds == c2g.Q_ADS).sum())} ads")
