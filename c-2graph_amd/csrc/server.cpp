// Query server: continuous batching of mixed SSSP / PPR / adsorption queries
// over one resident graph.
//
// The reference's collaboration loop (src/sssp/purn-perk/Main.cu:87-246) runs
// two GPU streams and one pthread CPU worker fed through a pinned ring
// (src/sssp/purn-perk/SSSP.cpp:18-84); it serves the same query repeatedly and
// moves work by queue length.  Here the same roles become lanes of a server:
//   * SSSP lane (stream A): a pool of query-major slots (the single-query
//     kernels run side by side) and one 64-lane group of the batched kernels
//     for the overflow.  A query is admitted into a free slot or lane between
//     rounds (its seed joins the next round's work) and retires as soon as
//     it has no queued work; its distances are copied out.  Requests that
//     arrive while others run therefore overlap with them on the device.
//   * PPR lane (stream B): the PPR requests queued when the lane frees up
//     (up to 64) run as one batched push, lane = query.
//   * adsorption lane (stream C): the adsorption requests queued when the lane
//     frees up run as one pass over their label sets side by side (up to 64
//     labels).  Both run concurrently with the SSSP rounds on the device.
//   * host lane (optional): the reference CPU worker's algorithms
//     (Bellman-Ford / forward push / adsorption Jacobi) on host threads, used
//     when the device is marked failed or a request cannot fit in HBM.
// Admission: every request reserves its device result buffer from a budget =
// free HBM at start (or opts.fake_free_hbm) minus opts.hbm_reserve; requests
// wait in FIFO order until the budget and a lane free up.  Submission beyond
// max_inflight outstanding requests is refused with C2G_EBUSY.
#include "graph.h"
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <unordered_map>

namespace c2g {

struct Request {
    int64_t id = 0;
    int kind = 0;
    int32_t src = 0;
    int status = 0;
    int lane = C2G_LANE_GPU;
    bool done = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t bytes = 0;
    std::unique_ptr<char[]> result;  // bytes long; from the server's buffer pool
};

}  // namespace c2g

struct c2g_server {
    c2g_graph* h = nullptr;
    c2g::Graph* g = nullptr;
    c2g_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> q_sssp, q_aux, q_ads, q_host;
    std::mutex pool_mu;  // taken after `mu` when both are held
    std::vector<std::unique_ptr<char[]>> pool[3];  // free result buffers per query kind
    std::unordered_map<int64_t, c2g::Request> reqs;
    int64_t next_id = 0;
    int64_t inflight = 0;
    int64_t budget = 0, used = 0;
    bool stop = false;
    std::thread t_sssp, t_aux, t_ads;
    std::vector<std::thread> t_host;
    hipStream_t s_sssp = nullptr, s_aux = nullptr, s_ads = nullptr;
    double t0 = 0;
    int error = 0;
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
    const int64_t V = s->g->V;
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

static void finish(c2g_server* s, Request& r, int status, int lane) {
    r.status = status;
    r.lane = lane;
    r.t_done = now_ms() - s->t0;
    r.done = true;
}

// SSSP lane.  Two continuously batched pools share one stream: the slot pool
// (query-major, the single-query kernels side by side) takes requests while
// it has a free slot; the 64-lane group (lane-major rows) takes the overflow.
// With few queries in flight a 256-byte lane row would carry mostly idle
// lanes, so light load stays in the slots; heavy load fills the lanes, where
// one row read serves many queries.
static void sssp_worker(c2g_server* s) {
    Graph* g = s->g;
    (void)hipSetDevice(g->device);
    const int nslots = s->o.sssp_slots < 0 ? 0 : std::min(64, s->o.sssp_slots == 0 ? 16 : s->o.sssp_slots);
    const int nlanes = std::min(64, std::max(1, s->o.max_batch));
    int rc = nslots > 0 ? qserve_init(g, s->s_sssp, nslots) : C2G_OK;
    bool lanes_ready = false;
    int64_t slot_req[64], lane_req[64];
    for (int i = 0; i < 64; ++i) slot_req[i] = lane_req[i] = -1;
    auto busy = [&](const int64_t* a, int n) {
        for (int i = 0; i < n; ++i)
            if (a[i] >= 0) return true;
        return false;
    };
    while (true) {
        std::vector<std::pair<int, int32_t>> admit_s, admit_l;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!busy(slot_req, nslots) && !busy(lane_req, nlanes))
                s->cv_work.wait(lk, [&] { return s->stop || !s->q_sssp.empty(); });
            if (s->stop) break;
            int si = 0, li = 0;
            while (!s->q_sssp.empty()) {
                while (si < nslots && slot_req[si] >= 0) ++si;
                while (li < nlanes && lane_req[li] >= 0) ++li;
                if (si >= nslots && li >= nlanes) break;
                Request& r = s->reqs[s->q_sssp.front()];
                if (s->used + r.bytes > s->budget) break;  // wait for memory
                s->q_sssp.pop_front();
                s->used += r.bytes;
                r.t_admit = now_ms() - s->t0;
                if (si < nslots) {
                    slot_req[si] = r.id;
                    admit_s.emplace_back(si, r.src);
                } else {
                    lane_req[li] = r.id;
                    admit_l.emplace_back(li, r.src);
                }
            }
        }
        const bool sbusy = busy(slot_req, nslots), lbusy = busy(lane_req, nlanes);
        if (!sbusy && !lbusy) {  // head request waits for HBM budget held by another lane
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        if (!admit_l.empty() && !lanes_ready && rc == C2G_OK) {
            rc = bserve_init(g, s->s_sssp);
            lanes_ready = rc == C2G_OK;
        }
        for (auto& a : admit_s)
            if (rc == C2G_OK) rc = qserve_admit(g, s->s_sssp, a.first, a.second);
        for (auto& a : admit_l)
            if (rc == C2G_OK) rc = bserve_admit(g, s->s_sssp, a.first, a.second);
        uint64_t swork = 0, lwork = 0;
        if (sbusy && rc == C2G_OK) rc = qserve_step(g, s->s_sssp, &swork);
        if (lbusy && rc == C2G_OK) rc = bserve_step(g, s->s_sssp, &lwork);
        // Retire queries with no queued work (all of them after an error).
        std::vector<int> done_s, lanes;
        std::vector<Request*> rs_s, rs_l;
        std::vector<int32_t*> outs;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int i = 0; i < nslots; ++i)
                if (slot_req[i] >= 0 && (rc != C2G_OK || !((swork >> i) & 1ull))) {
                    done_s.push_back(i);
                    rs_s.push_back(&s->reqs[slot_req[i]]);
                }
            for (int l = 0; l < nlanes; ++l)
                if (lane_req[l] >= 0 && (rc != C2G_OK || !((lwork >> l) & 1ull))) {
                    lanes.push_back(l);
                    rs_l.push_back(&s->reqs[lane_req[l]]);
                }
        }
        for (Request* r : rs_s) take_buffer(s, r);
        for (Request* r : rs_l) {
            take_buffer(s, r);
            outs.push_back((int32_t*)r->result.get());
        }
        for (size_t k = 0; k < done_s.size(); ++k)
            if (rc == C2G_OK) rc = qserve_extract(g, s->s_sssp, done_s[k], (int32_t*)rs_s[k]->result.get());
        if (!lanes.empty() && rc == C2G_OK) rc = bserve_extract(g, s->s_sssp, lanes.data(), (int)lanes.size(), outs.data());
        if (!done_s.empty() || !lanes.empty()) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int i : done_s) {
                Request& r = s->reqs[slot_req[i]];
                s->used -= r.bytes;
                finish(s, r, rc, C2G_LANE_GPU);
                slot_req[i] = -1;
            }
            for (int l : lanes) {
                Request& r = s->reqs[lane_req[l]];
                s->used -= r.bytes;
                finish(s, r, rc, C2G_LANE_GPU);
                lane_req[l] = -1;
            }
            s->cv_done.notify_all();
            s->cv_work.notify_all();  // memory freed: other lanes may admit
        }
        if (rc != C2G_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->error = rc;
        }
    }
}

// One device lane per kind (PPR on one stream, adsorption on another), so a
// short PPR query never waits behind a long adsorption run.
static void aux_worker(c2g_server* s, bool ads) {
    Graph* g = s->g;
    (void)hipSetDevice(g->device);
    std::deque<int64_t>& q = ads ? s->q_ads : s->q_aux;
    hipStream_t strm = ads ? s->s_ads : s->s_aux;
    // Requests queued together run as one pass: PPR as one batched push (lane =
    // query, up to 64), adsorption as one run over the queries' label sets side
    // by side (up to 64 labels in all).
    const size_t cap = ads ? (size_t)std::max(1, 64 / std::max(1, s->o.ads_labels))
                           : (size_t)std::min(64, std::max(1, s->o.max_batch));
    std::vector<std::vector<int32_t>> svs, sls;
    std::vector<float*> youts;
    std::vector<int32_t> srcs;
    std::vector<int64_t> ids;
    std::vector<float*> outs;
    while (true) {
        ids.clear();
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop) return true;
                if (q.empty()) return false;
                return s->used + s->reqs[q.front()].bytes <= s->budget;
            });
            if (s->stop) break;
            const double t_admit = now_ms() - s->t0;
            while (!q.empty() && ids.size() < cap && s->used + s->reqs[q.front()].bytes <= s->budget) {
                Request& r = s->reqs[q.front()];
                q.pop_front();
                s->used += r.bytes;
                r.t_admit = t_admit;
                ids.push_back(r.id);
            }
        }
        std::vector<Request*> rs;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (int64_t id : ids) rs.push_back(&s->reqs[id]);
        }
        for (Request* r : rs) take_buffer(s, r);
        int rc;
        if (ads) {
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
        } else if (rs.size() == 1) {
            rc = ppr_single(g, strm, rs[0]->src, s->o.alpha, s->o.eps, (float*)rs[0]->result.get(), nullptr, nullptr);
        } else {
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
        for (Request* r : rs) {
            s->used -= r->bytes;
            finish(s, *r, rc, C2G_LANE_GPU);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // memory freed: the SSSP lane may admit
    }
}

static void host_worker(c2g_server* s) {
    Graph* g = s->g;
    std::vector<int32_t> sv, sl;
    while (true) {
        int64_t id;
        Request* r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->q_host.empty(); });
            if (s->stop) break;
            id = s->q_host.front();
            s->q_host.pop_front();
            r = &s->reqs[id];
            r->t_admit = now_ms() - s->t0;
        }
        take_buffer(s, r);
        int rc = C2G_OK;
        if (r->kind == C2G_Q_SSSP) {
            int32_t* d = (int32_t*)r->result.get();
            for (int32_t v = 0; v < g->V; ++v) d[v] = INT32_MAX;
            d[r->src] = 0;
            std::vector<uint8_t> hid;
            if (!g->hvl.empty()) { hid.assign((size_t)g->V, 0); hid[r->src] = 1; }
            rc = host_bf(g, d, &r->src, 1, hid.empty() ? nullptr : hid.data(), nullptr);
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

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_server_start(c2g_graph* h, const c2g_server_opts* o, c2g_server** out) {
    if (!h || !out) return fail(C2G_EINVAL, "null argument");
    Graph* g = &h->g;
    auto* s = new c2g_server;
    s->h = h;
    s->g = g;
    if (o) s->o = *o;
    if (s->o.max_batch <= 0 || s->o.max_batch > 64) s->o.max_batch = 64;
    if (s->o.max_inflight <= 0) s->o.max_inflight = 1 << 20;
    if (s->o.alpha <= 0) s->o.alpha = 0.15f;
    if (s->o.eps <= 0) s->o.eps = 1e-7f;
    if (s->o.ads_labels <= 0) s->o.ads_labels = 16;
    if (s->o.ads_iters <= 0) s->o.ads_iters = 20;
    if (s->o.p_inj <= 0 && s->o.p_cont <= 0) { s->o.p_inj = 0.25f; s->o.p_cont = 0.70f; s->o.p_abnd = 0.05f; }
    if (s->o.ads_grid_per_cu < 0 || s->o.ads_grid_per_cu > 64) {
        delete s;
        return fail(C2G_EINVAL, "server: ads_grid_per_cu must be in [0,64]");
    }
    const bool device_ok = g->on_device && !s->o.device_failed;
    if (!device_ok && !s->o.host_lane) {
        delete s;
        return fail(C2G_EDEVICE, "server: no usable device and the host lane is disabled");
    }
    if (device_ok) {
        int rc;
        if (hipSetDevice(g->device) != hipSuccess) { delete s; return fail(C2G_EDEVICE, "hipSetDevice failed"); }
        size_t fr = 0, tot = 0;
        if (s->o.fake_free_hbm > 0) fr = (size_t)s->o.fake_free_hbm;
        else if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return fail(C2G_EDEVICE, "hipMemGetInfo failed"); }
        s->budget = (int64_t)fr - s->o.hbm_reserve;
        // The SSSP lane runs on the graph's own stream (idle while the server
        // owns the handle).  The HIP runtime maps a process's streams onto a
        // few hardware queues (4 by default, the null stream included); a
        // lane whose stream shares a queue with another lane waits behind
        // that lane's kernels in order.  With no extra SSSP stream, the
        // PPR and adsorption lanes each keep a queue of their own.
        s->s_sssp = g->stream;
        if (hipStreamCreateWithFlags(&s->s_aux, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&s->s_ads, hipStreamNonBlocking) != hipSuccess) {
            delete s;
            return fail(C2G_EDEVICE, "stream creation failed");
        }
        if (!g->d_adj && !g->d_adjf) { delete s; return fail(C2G_ENOTSUP, "graph has neither weight kind"); }
        (void)rc;
    }
    s->t0 = now_ms();
    if (device_ok) {
        if (g->d_adj) s->t_sssp = std::thread(sssp_worker, s);
        if (g->d_adjf) {
            s->t_aux = std::thread(aux_worker, s, false);
            s->t_ads = std::thread(aux_worker, s, true);
        }
    }
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
    Request r;
    r.id = s->next_id++;
    r.kind = kind;
    r.src = src;
    r.bytes = result_bytes(s, kind);
    r.t_submit = now_ms() - s->t0;
    const bool device_ok = g->on_device && !s->o.device_failed;
    const bool has_dev_lane = device_ok && (kind == C2G_Q_SSSP ? g->d_adj != nullptr : g->d_adjf != nullptr);
    const bool fits = r.bytes <= s->budget;
    if (!has_dev_lane || !fits) {
        if (!s->o.host_lane)
            return fail(C2G_EBUSY, "server: request needs %lld bytes of HBM, budget %lld", (long long)r.bytes,
                        (long long)s->budget);
        s->q_host.push_back(r.id);
    } else if (kind == C2G_Q_SSSP) {
        s->q_sssp.push_back(r.id);
    } else if (kind == C2G_Q_PPR) {
        s->q_aux.push_back(r.id);
    } else {
        s->q_ads.push_back(r.id);
    }
    s->reqs.emplace(r.id, std::move(r));
    s->inflight++;
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
    if (st < 0) return fail(st, "request %lld failed", (long long)id);
    return 1;
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
    if (s->t_aux.joinable()) s->t_aux.join();
    if (s->t_ads.joinable()) s->t_ads.join();
    for (auto& t : s->t_host) t.join();
    if (s->s_sssp) (void)hipStreamSynchronize(s->s_sssp);  // the graph's stream: not ours to destroy
    if (s->s_aux) { (void)hipStreamSynchronize(s->s_aux); (void)hipStreamDestroy(s->s_aux); }
    if (s->s_ads) { (void)hipStreamSynchronize(s->s_ads); (void)hipStreamDestroy(s->s_ads); }
    delete s;
    return C2G_OK;
}

}  // extern "C"
