// convert2binary: SNAP-style adjacency text -> origin/{vlist,elist,wlist,flist}.bin
// (utils/convert2binary.cpp:10-135).
//
// Input lines are "src dst1:dst2:..."; lines that do not parse as an integer
// and a token are skipped (:24-28).  V = largest id seen + 1 (:30-44); self
// loops are dropped (:67); every row is sorted and deduplicated (:76-79).
// Weights are (rand() % 100) + 1 drawn in CSR order from the C library's
// generator left at its default seed (:108), and flist = w / row sum (:111-113).
// The generator is the GNU C library's additive feedback generator (the
// TYPE_3 state of random(3): degree 31, separation 3, seeded by a
// multiplicative congruential sequence 16807 mod 2^31-1, 310 outputs
// discarded), restated here so that weights match the reference's on any host.
#include "common.h"
#include <cstring>
#include <cerrno>
#include <fstream>
#include <sstream>
#include <string>

namespace c2g {

int mkdirs(const std::string& d);
int write_vec(const std::string& p, const void* d, int64_t n);

struct GlibcRand {
    std::vector<int32_t> st;
    explicit GlibcRand(uint32_t seed) {
        st.resize(34);
        int32_t x = (int32_t)(seed ? seed : 1);
        st[0] = x;
        for (int k = 1; k < 31; ++k) {
            // 16807 * x mod (2^31 - 1) without overflow (Schrage's method)
            const int32_t hi = x / 127773, lo = x % 127773;
            int32_t w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            x = w;
            st[k] = x;
        }
        for (int k = 31; k < 34; ++k) st[k] = st[k - 31];
        buf.assign(st.begin(), st.begin() + 34);
        for (int k = 34; k < 344; ++k) next_raw();
    }
    std::vector<int32_t> buf;  // sliding window of the last 34 values
    uint32_t next_raw() {
        const size_t n = buf.size();
        const uint32_t v = (uint32_t)buf[n - 31] + (uint32_t)buf[n - 3];
        buf.push_back((int32_t)v);
        if (buf.size() > 4096) buf.erase(buf.begin(), buf.end() - 34);
        return v;
    }
    int next() { return (int)(next_raw() >> 1); }
};

}  // namespace c2g

using namespace c2g;

extern "C" int c2g_convert_adjlist(const char* txt, const char* ds) {
    if (!txt || !ds) return fail(C2G_EINVAL, "null argument");
    std::ifstream in(txt);
    if (!in.is_open()) return fail(C2G_EIO, "Error opening file: %s", txt);
    std::vector<std::pair<int32_t, int32_t>> edges;
    int64_t vmax = 0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        long long src;
        std::string part;
        if (!(iss >> src >> part)) continue;
        if (src < 0 || src > INT32_MAX - 1) return fail(C2G_EFORMAT, "vertex id %lld out of range", src);
        vmax = std::max<int64_t>(vmax, src);
        std::stringstream ds2(part);
        std::string tok;
        while (std::getline(ds2, tok, ':')) {
            char* end = nullptr;
            errno = 0;
            long long d = strtoll(tok.c_str(), &end, 10);
            if (end == tok.c_str() || errno || d < 0 || d > INT32_MAX - 1)
                return fail(C2G_EFORMAT, "bad destination '%s' in line: %s", tok.c_str(), line.c_str());
            vmax = std::max<int64_t>(vmax, d);
            if (d != src) edges.emplace_back((int32_t)src, (int32_t)d);
        }
    }
    const int64_t V = vmax + 1;
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    if ((int64_t)edges.size() > INT32_MAX) return fail(C2G_EFORMAT, "more than 2^31-1 edges");
    std::vector<int32_t> vl((size_t)V + 1, 0), el(edges.size()), wl(edges.size());
    std::vector<float> fl(edges.size());
    for (const auto& e : edges) vl[e.first + 1]++;
    for (int64_t v = 0; v < V; ++v) vl[v + 1] += vl[v];
    GlibcRand rnd(1);
    for (int64_t v = 0; v < V; ++v) {
        int sum = 0;
        for (int32_t j = vl[v]; j < vl[v + 1]; ++j) {
            el[j] = edges[j].second;
            wl[j] = (rnd.next() % 100) + 1;
            sum += wl[j];
        }
        for (int32_t j = vl[v]; j < vl[v + 1]; ++j) fl[j] = (float)(1.0 * wl[j] / sum);
    }
    printf("Max Vertex Number: %lld\nVertex Number: %lld\nEdge Number: %zu\n", (long long)vmax, (long long)V,
           edges.size());
    return c2g_write_dataset(ds, (int32_t)V, vl.data(), el.data(), wl.data(), fl.data());
}
