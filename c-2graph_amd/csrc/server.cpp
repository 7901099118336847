// Query server: continuous batching of mixed SSSP / PPR / adsorption queries
// over one resident graph.
//
// The reference's collaboration loop (src/sssp/purn-perk/Main.cu:87-246) runs
// two GPU streams and one pthread CPU worker fed through a pinned ring
// (src/sssp/purn-perk/SSSP.cpp:18-84); it serves the same query repeatedly and
// moves work by queue length (:128-145).  Here the same roles become lanes:
//   * SSSP (the graph's stream): a pool of query-major slots (the single-query
//     kernels run side by side) and one 64-lane group of the batched kernels
//     for the overflow.  A query joins a free slot or lane between rounds and
//     retires as soon as it has no queued work, so requests that arrive while
//     others run overlap with them on the device.
//   * PPR (stream B): the PPR requests queued when the lane frees up (up to
//     64) run as one batched push, lane = query.
//   * adsorption (stream C): the queued adsorption requests run as one pass
//     over their label sets side by side (up to 64 labels).
//   * host (optional): the reference CPU worker's algorithms (Bellman-Ford /
//     forward push / adsorption Jacobi) on host threads, for kinds without a
//     device lane and for the running requests of a device lane that failed.
//
// Admission by device memory: each device lane's workspace (slot pool, lane
// group, batched PPR arrays, adsorption iterates) is charged at start against
// the budget = free HBM (or opts.fake_free_hbm) minus opts.hbm_reserve, in the
// order slots, group, PPR, adsorption; the slot count is halved until the pool
// fits, and a lane that does not fit is not started.  A request whose kind has
// no device lane goes to the host lane, or is refused with C2G_EBUSY.
// Requests that do have a lane wait in FIFO order for a slot, lane or pass;
// beyond max_inflight outstanding requests submission is refused (C2G_EBUSY).
//
// Failure: a device lane whose step fails hands its running requests to the
// host lane (or finishes them with the error when the host lane is off) and is
// reset; a lane whose reset fails, or that fails twice in a row, is retired.
// Cancellation (c2g_server_cancel) and deadlines (opts.deadline_ms) finish a
// queued request at once and a running one at its lane's next step boundary,
// with C2G_ETIMEOUT.
//
// The device lanes sit behind a small backend interface: the HIP pools, or
// (opts.dry_device, tests) host computations, so admission, failover,
// cancellation and deadlines are exercised without a GPU.
#include "graph.h"
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <unordered_map>

namespace c2g {

enum { L_SLOTS = 0, L_GROUP = 1, L_PPR = 2, L_ADS = 3, L_N = 4 };

struct Request {
    int64_t id = 0;
    int kind = 0;
    int32_t src = 0;
    int status = 0;
    int lane = C2G_LANE_GPU;
    bool done = false;
    bool running = false;  // admitted to a lane
    bool cancel = false;   // finish with C2G_ETIMEOUT
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t bytes = 0;                 // host result buffer
    std::unique_ptr<char[]> result;    // bytes long; from the server's buffer pool
};

}  // namespace c2g

struct c2g_server {
    c2g_graph* h = nullptr;
    c2g::Graph* g = nullptr;
    c2g_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> q_sssp, q_ppr, q_ads, q_host;
    std::mutex pool_mu;  // taken after `mu` when both are held
    std::vector<std::unique_ptr<char[]>> pool[3];  // free result buffers per query kind
    std::unordered_map<int64_t, c2g::Request> reqs;
    int64_t next_id = 0;
    int64_t inflight = 0;
    int64_t budget = 0, charged = 0;
    int64_t lane_bytes[c2g::L_N] = {};
    bool lane_on[c2g::L_N] = {};  // started and not retired (written under mu)
    int nslots = 0, nlanes = 0;
    bool device = false;          // device lanes exist (a GPU, or dry_device)
    bool dry = false;
    bool stop = false;
    // failure injection (tests): touched only by each lane's own thread
    int steps[c2g::L_N] = {};
    bool injected[c2g::L_N] = {};
    std::thread t_sssp, t_ppr, t_ads;
    std::vector<std::thread> t_host;
    hipStream_t s_sssp = nullptr, s_aux = nullptr, s_ads = nullptr;
    double t0 = 0;
    int64_t n_submitted = 0, n_completed = 0, n_failed = 0, n_cancelled = 0, n_to_host = 0, n_lane_fail = 0;
};

namespace c2g {

// Result buffers are recycled per query kind: a fresh V-sized (or V*L-sized)
// buffer would be zero-filled or page-faulted on every request.  Neither call
// may hold `mu`'s critical section for the copy itself.
constexpr size_t kPoolMax = 64;
static void take_buffer(c2g_server* s, Request* r) {
    std::unique_ptr<char[]> b;
    {
        std::lock_guard<std::mutex> lk(s->pool_mu);
        auto& fl = s->pool[r->kind];
        if (!fl.empty()) {
            b = std::move(fl.back());
            fl.pop_back();
        }
    }
    if (!b) b.reset(new char[(size_t)r->bytes]);
    r->result = std::move(b);
}
static void give_back(c2g_server* s, int kind, std::unique_ptr<char[]> b) {
    if (!b) return;
    std::lock_guard<std::mutex> lk(s->pool_mu);
    if (s->pool[kind].size() < kPoolMax) s->pool[kind].push_back(std::move(b));
}

static int64_t result_bytes(const c2g_server* s, int kind) {
    const int64_t V = std::max<int64_t>(s->g->V, 1);
    if (kind == C2G_Q_SSSP) return 4 * V;
    if (kind == C2G_Q_PPR) return 4 * V;
    return 4 * V * s->o.ads_labels;
}

// Seed set `k` of adsorption queries: labels on a seeded 1% of vertices.
static void ads_seeds(const Graph* g, int32_t k, int L, std::vector<int32_t>& sv, std::vector<int32_t>& sl) {
    const int64_t n = std::max<int64_t>(1, g->V / 100);
    Stream rs((uint64_t)k, 11);
    sv.resize(n);
    sl.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t h = rs((uint64_t)i);
        sv[i] = (int32_t)((h >> 32) % (uint64_t)g->V);
        sl[i] = (int32_t)((uint32_t)h % (uint32_t)L);
    }
}

static double now_rel(const c2g_server* s) { return now_ms() - s->t0; }

static bool expired(const c2g_server* s, const Request& r, double now) {
    return s->o.deadline_ms > 0 && now - r.t_submit > (double)s->o.deadline_ms;
}

// Completes `r` (caller holds mu).  A cancelled or expired request reports
// C2G_ETIMEOUT whatever its lane produced.
static void finish(c2g_server* s, Request& r, int status, int lane) {
    const double now = now_rel(s);
    if (status == C2G_OK && (r.cancel || expired(s, r, now))) status = C2G_ETIMEOUT;
    r.status = status;
    r.lane = lane;
    r.t_done = now;
    r.done = true;
    r.running = false;
    if (status == C2G_OK) ++s->n_completed;
    else if (status == C2G_ETIMEOUT) ++s->n_cancelled;
    else ++s->n_failed;
}

// Hands `r` to the host lane, or finishes it with `rc` when that lane is off
// (caller holds mu).
static void to_host_or_fail(c2g_server* s, Request& r, int rc) {
    r.running = false;
    if (s->o.host_lane) {
        s->q_host.push_back(r.id);
        ++s->n_to_host;
    } else {
        finish(s, r, rc, C2G_LANE_GPU);
    }
}

// Drops cancelled and expired requests from a queue (caller holds mu).
static void sweep(c2g_server* s, std::deque<int64_t>& q, double now) {
    for (size_t i = 0; i < q.size();) {
        Request& r = s->reqs[q[i]];
        if (r.cancel || expired(s, r, now)) {
            finish(s, r, C2G_ETIMEOUT, C2G_LANE_GPU);
            q.erase(q.begin() + (long)i);
        } else {
            ++i;
        }
    }
}

// Test hook: lane `k` reports a device error once, after fail_after steps.
static int inject(c2g_server* s, int k) {
    if (!(s->o.fail_inject & (1 << k)) || s->injected[k]) return C2G_OK;
    if (s->steps[k]++ < s->o.fail_after) return C2G_OK;
    s->injected[k] = true;
    return fail(C2G_EDEVICE, "injected failure of server lane %d", k);
}

// Host lane's single SSSP query: Bellman-Ford from {s:0}, frontier {s}
// (src/sssp/base-perk/SSSP.cpp:88-119), the source's hidden list relaxed on
// pruned graphs.
static int host_sssp_into(const Graph* g, int32_t src, int32_t* d) {
    for (int32_t v = 0; v < g->V; ++v) d[v] = INT32_MAX;
    d[src] = 0;
    std::vector<uint8_t> hid;
    if (!g->hvl.empty()) {
        hid.assign((size_t)g->V, 0);
        hid[src] = 1;
    }
    return host_bf(g, d, &src, 1, hid.empty() ? nullptr : hid.data(), nullptr);
}

// --------------------------------------------------------- SSSP backends
// One pool of SSSP query positions (slots, or lanes of the group).
struct SsspPool {
    virtual ~SsspPool() {}
    virtual int init() = 0;  // (re)start empty
    virtual int admit(int pos, int32_t src) = 0;
    virtual int step(uint64_t* work) = 0;  // positions with queued work after one round
    // Copies out (or, for a null out, drops) the queries at `pos`.
    virtual int extract(const std::vector<int>& pos, const std::vector<int32_t*>& outs) = 0;
};

struct SlotPoolDev : SsspPool {
    Graph* g;
    hipStream_t st;
    int n;
    SlotPoolDev(Graph* g_, hipStream_t s_, int n_) : g(g_), st(s_), n(n_) {}
    int init() override { return qserve_init(g, st, n); }
    int admit(int pos, int32_t src) override { return qserve_admit(g, st, pos, src); }
    int step(uint64_t* w) override { return qserve_step(g, st, w); }
    int extract(const std::vector<int>& pos, const std::vector<int32_t*>& outs) override {
        // a dropped slot keeps its state until qserve_admit resets it
        for (size_t k = 0; k < pos.size(); ++k)
            if (outs[k]) {
                int rc = qserve_extract(g, st, pos[k], outs[k]);
                if (rc) return rc;
            }
        return C2G_OK;
    }
};

struct GroupDev : SsspPool {
    Graph* g;
    hipStream_t st;
    GroupDev(Graph* g_, hipStream_t s_) : g(g_), st(s_) {}
    int init() override { return bserve_init(g, st); }
    int admit(int pos, int32_t src) override { return bserve_admit(g, st, pos, src); }
    int step(uint64_t* w) override { return bserve_step(g, st, w); }
    int extract(const std::vector<int>& pos, const std::vector<int32_t*>& outs) override {
        // extraction also resets the lanes' columns (null outs included)
        return bserve_extract(g, st, pos.data(), (int)pos.size(), outs.data());
    }
};

// Host computations in place of kernels: a query completes in one step.
struct DryPool : SsspPool {
    const Graph* g;
    int ms;
    std::vector<int32_t> src;
    std::vector<std::vector<int32_t>> dist;
    DryPool(const Graph* g_, int ms_) : g(g_), ms(ms_), src(64, -1), dist(64) {}
    int init() override {
        std::fill(src.begin(), src.end(), -1);
        return C2G_OK;
    }
    int admit(int pos, int32_t s) override {
        src[pos] = s;
        dist[pos].clear();
        return C2G_OK;
    }
    int step(uint64_t* w) override {
        if (ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(ms));
        for (int p = 0; p < 64; ++p)
            if (src[p] >= 0 && dist[p].empty()) {
                dist[p].resize((size_t)g->V);
                int rc = host_sssp_into(g, src[p], dist[p].data());
                if (rc) return rc;
            }
        *w = 0;
        return C2G_OK;
    }
    int extract(const std::vector<int>& pos, const std::vector<int32_t*>& outs) override {
        for (size_t k = 0; k < pos.size(); ++k) {
            if (outs[k]) memcpy(outs[k], dist[pos[k]].data(), sizeof(int32_t) * g->V);
            src[pos[k]] = -1;
        }
        return C2G_OK;
    }
};

// SSSP lane: the slot pool takes requests while it has a free slot, the
// 64-lane group takes the overflow.  With few queries in flight a 256-byte
// lane row would carry mostly idle lanes, so light load stays in the slots;
// heavy load fills the lanes, where one row read serves many queries.
static void sssp_worker(c2g_server* s) {
    Graph* g = s->g;
    if (!s->dry) (void)hipSetDevice(g->device);
    const int cap[2] = {s->nslots, s->nlanes};
    const int lane_id[2] = {L_SLOTS, L_GROUP};
    std::unique_ptr<SsspPool> pool[2];
    if (s->dry) {
        pool[0].reset(new DryPool(g, s->o.dry_step_ms));
        pool[1].reset(new DryPool(g, s->o.dry_step_ms));
    } else {
        pool[0].reset(new SlotPoolDev(g, s->s_sssp, std::max(1, s->nslots)));
        pool[1].reset(new GroupDev(g, s->s_sssp));
    }
    bool ready[2] = {false, false};
    int fails[2] = {0, 0};  // consecutive failures per pool
    int64_t req[2][64];
    for (auto& a : req)
        for (auto& x : a) x = -1;
    auto busy = [&](int k) {
        for (int i = 0; i < cap[k]; ++i)
            if (req[k][i] >= 0) return true;
        return false;
    };
    // A step, admission or extraction of pool k failed: its running requests
    // move to the host lane (or fail); the pool is re-initialised before its
    // next use, and retired after two failures in a row.
    auto pool_failed = [&](int k, int rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->n_lane_fail;
        for (int i = 0; i < cap[k]; ++i)
            if (req[k][i] >= 0) {
                to_host_or_fail(s, s->reqs[req[k][i]], rc);
                req[k][i] = -1;
            }
        ready[k] = false;
        if (++fails[k] >= 2) s->lane_on[lane_id[k]] = false;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    };
    while (true) {
        std::vector<std::pair<int, int32_t>> adm[2];
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!busy(0) && !busy(1)) s->cv_work.wait(lk, [&] { return s->stop || !s->q_sssp.empty(); });
            if (s->stop) break;
            const double now = now_rel(s);
            sweep(s, s->q_sssp, now);
            if (!s->lane_on[L_SLOTS] && !s->lane_on[L_GROUP]) {  // both pools retired
                while (!s->q_sssp.empty()) {
                    Request& r = s->reqs[s->q_sssp.front()];
                    s->q_sssp.pop_front();
                    to_host_or_fail(s, r, C2G_EDEVICE);
                }
                s->cv_done.notify_all();
                s->cv_work.notify_all();
                if (!busy(0) && !busy(1)) break;
            }
            for (int k = 0; k < 2; ++k) {
                if (!s->lane_on[lane_id[k]]) continue;
                for (int i = 0; i < cap[k] && !s->q_sssp.empty(); ++i) {
                    if (req[k][i] >= 0) continue;
                    Request& r = s->reqs[s->q_sssp.front()];
                    s->q_sssp.pop_front();
                    r.running = true;
                    r.t_admit = now;
                    req[k][i] = r.id;
                    adm[k].emplace_back(i, r.src);
                }
            }
        }
        uint64_t work[2] = {0, 0};
        for (int k = 0; k < 2; ++k) {
            if (!busy(k)) continue;
            int rc = C2G_OK;
            if (!ready[k]) {
                rc = pool[k]->init();
                ready[k] = rc == C2G_OK;
            }
            for (size_t a = 0; a < adm[k].size() && rc == C2G_OK; ++a) rc = pool[k]->admit(adm[k][a].first, adm[k][a].second);
            if (rc == C2G_OK) rc = inject(s, lane_id[k]);
            if (rc == C2G_OK) rc = pool[k]->step(&work[k]);
            if (rc) pool_failed(k, rc);
        }
        // Retire queries with no queued work, and cancelled or expired ones.
        for (int k = 0; k < 2; ++k) {
            std::vector<int> pos;
            std::vector<int32_t*> outs;
            std::vector<Request*> rs;
            std::vector<bool> drop;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                const double now = now_rel(s);
                for (int i = 0; i < cap[k]; ++i) {
                    if (req[k][i] < 0) continue;
                    Request& r = s->reqs[req[k][i]];
                    const bool dr = r.cancel || expired(s, r, now);
                    if (!dr && ((work[k] >> i) & 1ull)) continue;
                    pos.push_back(i);
                    rs.push_back(&r);
                    drop.push_back(dr);
                }
            }
            if (pos.empty()) continue;
            for (size_t j = 0; j < rs.size(); ++j) {
                if (!drop[j]) take_buffer(s, rs[j]);
                outs.push_back(drop[j] ? nullptr : (int32_t*)rs[j]->result.get());
            }
            const int rc = pool[k]->extract(pos, outs);
            if (rc) {
                pool_failed(k, rc);
                continue;
            }
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t j = 0; j < pos.size(); ++j) {
                finish(s, *rs[j], drop[j] ? C2G_ETIMEOUT : C2G_OK, C2G_LANE_GPU);
                req[k][pos[j]] = -1;
            }
            fails[k] = 0;
            s->cv_done.notify_all();
            s->cv_work.notify_all();
        }
    }
}

// One device lane per kind (PPR on one stream, adsorption on another), so a
// short PPR query never waits behind a long adsorption run.
static void aux_worker(c2g_server* s, bool ads) {
    Graph* g = s->g;
    if (!s->dry) (void)hipSetDevice(g->device);
    const int lane = ads ? L_ADS : L_PPR;
    std::deque<int64_t>& q = ads ? s->q_ads : s->q_ppr;
    hipStream_t strm = ads ? s->s_ads : s->s_aux;
    // Requests queued together run as one pass: PPR as one batched push (lane =
    // query, up to 64), adsorption as one run over the queries' label sets side
    // by side (up to 64 labels in all).
    const size_t cap = ads ? (size_t)std::max(1, 64 / std::max(1, s->o.ads_labels))
                           : (size_t)std::min(64, std::max(1, s->o.max_batch));
    std::vector<std::vector<int32_t>> svs, sls;
    std::vector<float*> youts;
    std::vector<int32_t> srcs;
    std::vector<float*> outs;
    std::vector<Request*> rs;
    int fails = 0;
    while (true) {
        rs.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !q.empty(); });
            if (s->stop) break;
            const double now = now_rel(s);
            sweep(s, q, now);
            if (!s->lane_on[lane]) {  // retired: everything queued goes elsewhere
                while (!q.empty()) {
                    Request& r = s->reqs[q.front()];
                    q.pop_front();
                    to_host_or_fail(s, r, C2G_EDEVICE);
                }
                s->cv_done.notify_all();
                s->cv_work.notify_all();
                break;
            }
            while (!q.empty() && rs.size() < cap) {
                Request& r = s->reqs[q.front()];
                q.pop_front();
                r.running = true;
                r.t_admit = now;
                rs.push_back(&r);
            }
            s->cv_done.notify_all();
        }
        if (rs.empty()) continue;
        for (Request* r : rs) take_buffer(s, r);
        int rc = inject(s, lane);
        if (rc == C2G_OK && s->dry) {
            if (s->o.dry_step_ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(s->o.dry_step_ms));
            std::vector<int32_t> sv, sl;
            for (Request* r : rs) {
                if (ads) {
                    ads_seeds(g, r->src, s->o.ads_labels, sv, sl);
                    rc = host_ads(g, sv.data(), sl.data(), (int32_t)sv.size(), s->o.ads_labels, s->o.p_inj,
                                  s->o.p_cont, s->o.ads_iters, (float*)r->result.get());
                } else {
                    rc = c2g_host_ppr(s->h, r->src, s->o.alpha, s->o.eps, (float*)r->result.get(), nullptr, nullptr);
                }
                if (rc) break;
            }
        } else if (rc == C2G_OK && ads) {
            const int k = (int)rs.size();
            svs.resize(k);
            sls.resize(k);
            youts.clear();
            for (int j = 0; j < k; ++j) {
                ads_seeds(g, rs[j]->src, s->o.ads_labels, svs[j], sls[j]);
                youts.push_back((float*)rs[j]->result.get());
            }
            rc = ads_run_batch(g, strm, k, svs.data(), sls.data(), s->o.ads_labels, s->o.p_inj, s->o.p_cont,
                               s->o.p_abnd, s->o.ads_iters, youts.data(), s->o.ads_grid_per_cu);
        } else if (rc == C2G_OK && rs.size() == 1) {
            rc = ppr_single(g, strm, rs[0]->src, s->o.alpha, s->o.eps, (float*)rs[0]->result.get(), nullptr, nullptr);
        } else if (rc == C2G_OK) {
            srcs.clear();
            outs.clear();
            for (Request* r : rs) {
                srcs.push_back(r->src);
                outs.push_back((float*)r->result.get());
            }
            rc = ppr_batch_to(g, strm, srcs.data(), (int32_t)srcs.size(), s->o.alpha, s->o.eps, outs.data(),
                              nullptr, nullptr);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) {
            // the lane's state is rebuilt by its next call; two failures in a
            // row retire it
            ++s->n_lane_fail;
            for (Request* r : rs) to_host_or_fail(s, *r, rc);
            if (++fails >= 2) s->lane_on[lane] = false;
        } else {
            fails = 0;
            for (Request* r : rs) finish(s, *r, C2G_OK, C2G_LANE_GPU);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

static void host_worker(c2g_server* s) {
    Graph* g = s->g;
    std::vector<int32_t> sv, sl;
    while (true) {
        Request* r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->q_host.empty(); });
            if (s->stop) break;
            r = &s->reqs[s->q_host.front()];
            s->q_host.pop_front();
            if (r->cancel || expired(s, *r, now_rel(s))) {
                finish(s, *r, C2G_ETIMEOUT, C2G_LANE_HOST);
                s->cv_done.notify_all();
                continue;
            }
            r->running = true;
            r->t_admit = now_rel(s);
        }
        take_buffer(s, r);
        int rc = C2G_OK;
        if (r->kind == C2G_Q_SSSP) {
            rc = host_sssp_into(g, r->src, (int32_t*)r->result.get());
        } else if (r->kind == C2G_Q_PPR) {
            rc = c2g_host_ppr(s->h, r->src, s->o.alpha, s->o.eps, (float*)r->result.get(), nullptr, nullptr);
        } else {
            ads_seeds(g, r->src, s->o.ads_labels, sv, sl);
            rc = host_ads(g, sv.data(), sl.data(), (int32_t)sv.size(), s->o.ads_labels, s->o.p_inj, s->o.p_cont,
                          s->o.ads_iters, (float*)r->result.get());
        }
        std::lock_guard<std::mutex> lk(s->mu);
        finish(s, *r, rc, C2G_LANE_HOST);
        s->cv_done.notify_all();
    }
}

// Charges the device lanes' workspaces against the budget, in order (caller
// owns `s`, no lanes running yet).
static void size_lanes(c2g_server* s, bool has_w, bool has_f) {
    Graph* g = s->g;
    auto charge = [&](int lane, int64_t bytes) {
        if (s->charged + bytes > s->budget) return false;
        s->charged += bytes;
        s->lane_bytes[lane] = bytes;
        s->lane_on[lane] = true;
        return true;
    };
    const int want = s->o.sssp_slots < 0 ? 0 : std::min(64, s->o.sssp_slots == 0 ? 16 : s->o.sssp_slots);
    if (has_w)
        for (int n = want; n >= 1; n /= 2)
            if (charge(L_SLOTS, qserve_ws_bytes(g, n))) {
                s->nslots = n;
                break;
            }
    if (has_w && charge(L_GROUP, bserve_ws_bytes(g))) s->nlanes = std::min(64, std::max(1, s->o.max_batch));
    if (has_f) charge(L_PPR, ppr_lane_ws_bytes(g));
    if (has_f) charge(L_ADS, ads_lane_ws_bytes(g, (64 / std::max(1, s->o.ads_labels)) * s->o.ads_labels));
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_server_start(c2g_graph* h, const c2g_server_opts* o, c2g_server** out) {
    if (!h || !out) return fail(C2G_EINVAL, "null argument");
    Graph* g = &h->g;
    c2g_server_opts opt{};
    if (o) opt = *o;
    if (opt.max_batch <= 0 || opt.max_batch > 64) opt.max_batch = 64;
    if (opt.max_inflight <= 0) opt.max_inflight = 1 << 20;
    if (opt.alpha <= 0) opt.alpha = 0.15f;
    if (opt.eps <= 0) opt.eps = 1e-7f;
    if (opt.ads_labels <= 0) opt.ads_labels = 16;
    if (opt.ads_labels > 64) return fail(C2G_EINVAL, "server: ads_labels must be in [1,64]");
    if (opt.ads_iters <= 0) opt.ads_iters = 20;
    if (opt.p_inj <= 0 && opt.p_cont <= 0) { opt.p_inj = 0.25f; opt.p_cont = 0.70f; opt.p_abnd = 0.05f; }
    if (opt.ads_grid_per_cu < 0 || opt.ads_grid_per_cu > 64)
        return fail(C2G_EINVAL, "server: ads_grid_per_cu must be in [0,64]");
    if (opt.deadline_ms < 0 || opt.dry_step_ms < 0 || opt.fail_after < 0)
        return fail(C2G_EINVAL, "server: deadline_ms, dry_step_ms and fail_after must be >= 0");
    if (opt.dry_device && opt.fake_free_hbm <= 0)
        return fail(C2G_EINVAL, "server: dry_device needs fake_free_hbm as its HBM budget");
    const bool device_ok = (g->on_device || opt.dry_device) && !opt.device_failed;
    if (!device_ok && !opt.host_lane) return fail(C2G_EDEVICE, "server: no usable device and the host lane is disabled");
    int expect = 0;
    if (!g->serving.compare_exchange_strong(expect, 1)) return fail(C2G_EBUSY, "server: the handle already has a server");
    // A direct call that is still running finishes before the lanes start.
    { std::lock_guard<std::mutex> lk(g->mu); }
    auto* s = new c2g_server;
    s->h = h;
    s->g = g;
    s->o = opt;
    s->dry = opt.dry_device != 0;
    s->device = device_ok;
    auto bail = [&](int code, const char* msg) {
        if (s->s_aux) (void)hipStreamDestroy(s->s_aux);
        if (s->s_ads) (void)hipStreamDestroy(s->s_ads);
        delete s;
        g->serving = 0;
        return fail(code, "server: %s", msg);
    };
    if (device_ok) {
        size_t fr = 0, tot = 0;
        if (!s->dry && hipSetDevice(g->device) != hipSuccess) return bail(C2G_EDEVICE, "hipSetDevice failed");
        if (opt.fake_free_hbm > 0) fr = (size_t)opt.fake_free_hbm;
        else if (hipMemGetInfo(&fr, &tot) != hipSuccess) return bail(C2G_EDEVICE, "hipMemGetInfo failed");
        s->budget = (int64_t)fr - opt.hbm_reserve;
        const bool has_w = s->dry ? g->has_w() : g->d_adj != nullptr;
        const bool has_f = s->dry ? g->has_f() : g->d_adjf != nullptr;
        if (!has_w && !has_f) return bail(C2G_ENOTSUP, "graph has neither weight kind");
        size_lanes(s, has_w, has_f);
        if (!s->dry) {
            // The SSSP lane runs on the graph's own stream (idle while the
            // server owns the handle).  The HIP runtime maps a process's
            // streams onto a few hardware queues (4 by default, the null
            // stream included); a lane whose stream shares a queue with
            // another lane waits behind that lane's kernels in order.  With
            // no extra SSSP stream, the PPR and adsorption lanes each keep a
            // queue of their own.
            s->s_sssp = g->stream;
            if (hipStreamCreateWithFlags(&s->s_aux, hipStreamNonBlocking) != hipSuccess ||
                hipStreamCreateWithFlags(&s->s_ads, hipStreamNonBlocking) != hipSuccess)
                return bail(C2G_EDEVICE, "stream creation failed");
        }
    }
    s->t0 = now_ms();
    if (s->lane_on[L_SLOTS] || s->lane_on[L_GROUP]) s->t_sssp = std::thread(sssp_worker, s);
    if (s->lane_on[L_PPR]) s->t_ppr = std::thread(aux_worker, s, false);
    if (s->lane_on[L_ADS]) s->t_ads = std::thread(aux_worker, s, true);
    if (s->o.host_lane)
        for (int i = 0; i < 2; ++i) s->t_host.emplace_back(host_worker, s);
    *out = s;
    return C2G_OK;
}

int64_t c2g_server_submit(c2g_server* s, int32_t kind, int32_t src) {
    if (!s) return fail(C2G_EINVAL, "null server");
    if (kind < C2G_Q_SSSP || kind > C2G_Q_ADS) return fail(C2G_EINVAL, "unknown query kind %d", kind);
    Graph* g = s->g;
    if (kind != C2G_Q_ADS && (src < 0 || src >= g->V)) return fail(C2G_EINVAL, "source %d out of range", src);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(C2G_EINVAL, "server stopped");
    if (s->inflight >= s->o.max_inflight) return fail(C2G_EBUSY, "server: %lld requests in flight", (long long)s->inflight);
    const bool dev = s->device && (kind == C2G_Q_SSSP ? (s->lane_on[L_SLOTS] || s->lane_on[L_GROUP])
                                                      : s->lane_on[kind == C2G_Q_PPR ? L_PPR : L_ADS]);
    if (!dev && !s->o.host_lane)
        return fail(C2G_EBUSY, "server: no device lane serves query kind %d (HBM budget %lld bytes, %lld charged) and "
                               "the host lane is off", kind, (long long)s->budget, (long long)s->charged);
    Request r;
    r.id = s->next_id++;
    r.kind = kind;
    r.src = src;
    r.bytes = result_bytes(s, kind);
    r.t_submit = now_rel(s);
    if (!dev) {
        s->q_host.push_back(r.id);
        ++s->n_to_host;
    } else if (kind == C2G_Q_SSSP) {
        s->q_sssp.push_back(r.id);
    } else if (kind == C2G_Q_PPR) {
        s->q_ppr.push_back(r.id);
    } else {
        s->q_ads.push_back(r.id);
    }
    s->reqs.emplace(r.id, std::move(r));
    s->inflight++;
    s->n_submitted++;
    s->cv_work.notify_all();
    return s->next_id - 1;
}

int c2g_server_poll(c2g_server* s, int64_t id, int32_t wait_ms, void* out, c2g_result_info* info) {
    if (!s) return fail(C2G_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(C2G_EINVAL, "unknown request id %lld", (long long)id);
    auto ready = [&] { return s->reqs[id].done; };
    if (!ready()) {
        if (wait_ms == 0) return 0;
        if (wait_ms < 0) s->cv_done.wait(lk, ready);
        else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready)) return 0;
    }
    Request& r = s->reqs[id];
    if (info) {
        info->id = r.id;
        info->kind = r.kind;
        info->status = r.status;
        info->lane = r.lane;
        info->t_submit_ms = r.t_submit;
        info->t_admit_ms = r.t_admit;
        info->t_done_ms = r.t_done;
    }
    const int st = r.status, kind = r.kind;
    const size_t n = (size_t)r.bytes;
    std::unique_ptr<char[]> buf = std::move(r.result);
    s->reqs.erase(id);
    s->inflight--;
    lk.unlock();  // the copy runs without blocking the lanes
    if (out && st == C2G_OK && buf) memcpy(out, buf.get(), n);
    give_back(s, kind, std::move(buf));
    if (st == C2G_ETIMEOUT) return fail(st, "request %lld was cancelled or passed its deadline", (long long)id);
    if (st < 0) return fail(st, "request %lld failed", (long long)id);
    return 1;
}

int c2g_server_cancel(c2g_server* s, int64_t id) {
    if (!s) return fail(C2G_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(C2G_EINVAL, "unknown request id %lld", (long long)id);
    Request& r = it->second;
    if (r.done) return 0;
    r.cancel = true;
    if (!r.running) {  // still queued: take it out now
        for (auto* q : {&s->q_sssp, &s->q_ppr, &s->q_ads, &s->q_host})
            for (auto qi = q->begin(); qi != q->end(); ++qi)
                if (*qi == id) {
                    q->erase(qi);
                    finish(s, r, C2G_ETIMEOUT, C2G_LANE_GPU);
                    s->cv_done.notify_all();
                    return 1;
                }
    }
    return 1;  // running: its lane finishes it at the next step boundary
}

int c2g_server_get_stats(c2g_server* s, c2g_server_stats* out) {
    if (!s || !out) return fail(C2G_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof *out);
    out->budget = s->budget;
    out->charged = s->charged;
    for (int k = 0; k < L_N; ++k) {
        out->lane_bytes[k] = s->lane_on[k] ? s->lane_bytes[k] : 0;
        if (s->lane_on[k]) out->lanes_enabled |= 1 << k;
    }
    out->sssp_slots = s->lane_on[L_SLOTS] ? s->nslots : 0;
    out->submitted = s->n_submitted;
    out->completed = s->n_completed;
    out->failed = s->n_failed;
    out->cancelled = s->n_cancelled;
    out->to_host = s->n_to_host;
    out->lane_failures = s->n_lane_fail;
    out->queued = (int32_t)(s->q_sssp.size() + s->q_ppr.size() + s->q_ads.size() + s->q_host.size());
    int32_t run = 0;
    for (auto& kv : s->reqs) run += kv.second.running && !kv.second.done;
    out->running = run;
    return C2G_OK;
}

int c2g_server_stop(c2g_server* s) {
    if (!s) return C2G_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->t_sssp.joinable()) s->t_sssp.join();
    if (s->t_ppr.joinable()) s->t_ppr.join();
    if (s->t_ads.joinable()) s->t_ads.join();
    for (auto& t : s->t_host) t.join();
    Graph* g = s->g;
    if (!s->dry && s->device) {
        if (s->s_sssp) (void)hipStreamSynchronize(s->s_sssp);  // the graph's stream: not ours to destroy
        if (s->s_aux) { (void)hipStreamSynchronize(s->s_aux); (void)hipStreamDestroy(s->s_aux); }
        if (s->s_ads) { (void)hipStreamSynchronize(s->s_ads); (void)hipStreamDestroy(s->s_ads); }
        // The slot pool exists only for the server; the next server sizes its own.
        qserve_release(g);
        (void)hipGetLastError();
    }
    delete s;
    g->serving = 0;
    return C2G_OK;
}

}  // extern "C"
