// Shared helpers of the command-line tools: positional arguments as in the
// reference mains plus optional --key value flags.
#pragma once
#include <c2g.h>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

inline int die(int rc) {
    fprintf(stderr, "error %d: %s\n", rc, c2g_last_error());
    return 1;
}

struct Args {
    std::vector<std::string> pos;
    std::map<std::string, std::string> opt;
    bool has(const std::string& k) const { return opt.count(k) != 0; }
    std::string str(const std::string& k, const std::string& d) const {
        auto it = opt.find(k);
        return it == opt.end() ? d : it->second;
    }
    long num(const std::string& k, long d) const { return has(k) ? atol(opt.at(k).c_str()) : d; }
    double real(const std::string& k, double d) const { return has(k) ? atof(opt.at(k).c_str()) : d; }
};

// Flags listed in `bare` take no value.
inline Args parse_args(int argc, char** argv, const std::vector<std::string>& bare = {}) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        if (s.rfind("--", 0) == 0) {
            std::string k = s.substr(2);
            bool is_bare = false;
            for (auto& b : bare) is_bare |= b == k;
            if (is_bare || i + 1 >= argc) a.opt[k] = "1";
            else a.opt[k] = argv[++i];
        } else {
            a.pos.push_back(s);
        }
    }
    return a;
}

inline double wall_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

inline c2g_graph* open_graph(const std::string& ds, int variant, bool flist, int device) {
    c2g_opts o{};
    o.device = device;
    o.load_flist = flist ? 1 : 0;
    o.verbose = 1;
    c2g_graph* g = nullptr;
    int rc = c2g_open(ds.c_str(), variant, &o, &g);
    if (rc) { die(rc); return nullptr; }
    return g;
}

// Writes an int32/float32 array if --out was given.
inline int maybe_write(const Args& a, const void* buf, int64_t n, bool is_float) {
    if (!a.has("out")) return 0;
    int rc = is_float ? c2g_write_f32(a.str("out", "").c_str(), (const float*)buf, n)
                      : c2g_write_i32(a.str("out", "").c_str(), (const int32_t*)buf, n);
    return rc ? die(rc) : 0;
}
