// Binary CSR format: headerless little-endian int32 / float32 arrays.
//
// Reader semantics follow read_binary2vector<T> (core/io.hpp:25-40): the
// element count is filesize / 4.  Unlike the reference, a missing file is an
// error code, not exit(0) (core/io.hpp:27-30).
#include "common.h"
#include <sys/stat.h>
#include <cerrno>
#include <cstring>
#include <mutex>

namespace c2g {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
const char* get_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int host_threads() {
    static int n = [] {
        const char* e = getenv("C2G_THREADS");
        int v = e ? atoi(e) : 0;
        if (v <= 0) {
            v = (int)std::thread::hardware_concurrency();
            if (v <= 0) v = 1;
            v = std::min(v, 16);
        }
        return v;
    }();
    return n;
}

static int64_t read_raw(const char* path, void* buf, int64_t cap) {
    if (!path) return fail(C2G_EINVAL, "null path");
    struct stat st;
    if (stat(path, &st) != 0) return fail(C2G_EIO, "file:%s not exist.", path);
    int64_t cnt = (int64_t)st.st_size / 4;
    if (!buf) return cnt;
    if (cap < cnt) return fail(C2G_EINVAL, "%s: buffer holds %lld of %lld elements", path,
                               (long long)cap, (long long)cnt);
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(C2G_EIO, "%s: %s", path, strerror(errno));
    size_t got = fread(buf, 4, (size_t)cnt, fp);
    fclose(fp);
    if ((int64_t)got != cnt) return fail(C2G_EIO, "%s: short read", path);
    return cnt;
}

static int write_raw(const char* path, const void* buf, int64_t n) {
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(C2G_EIO, "%s: %s", path, strerror(errno));
    size_t put = n ? fwrite(buf, 4, (size_t)n, fp) : 0;
    int rc = fclose(fp);
    if ((int64_t)put != n || rc != 0) return fail(C2G_EIO, "%s: short write", path);
    return C2G_OK;
}

int mkdirs(const std::string& d) {
    std::string cur;
    for (size_t i = 0; i < d.size(); ++i) {
        cur.push_back(d[i]);
        if (d[i] == '/' || i + 1 == d.size()) {
            if (cur.size() > 1 || cur != "/") {
                if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                    return fail(C2G_EIO, "mkdir %s: %s", cur.c_str(), strerror(errno));
            }
        }
    }
    return C2G_OK;
}

bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// Reads a whole int32/float32 file into a vector; returns element count or error.
int64_t read_vec_i32(const std::string& p, std::vector<int32_t>& v) {
    int64_t n = read_raw(p.c_str(), nullptr, 0);
    if (n < 0) return n;
    v.resize((size_t)n);
    return n ? read_raw(p.c_str(), v.data(), n) : 0;
}
int64_t read_vec_f32(const std::string& p, std::vector<float>& v) {
    int64_t n = read_raw(p.c_str(), nullptr, 0);
    if (n < 0) return n;
    v.resize((size_t)n);
    return n ? read_raw(p.c_str(), v.data(), n) : 0;
}
int write_vec(const std::string& p, const void* d, int64_t n) { return write_raw(p.c_str(), d, n); }

}  // namespace c2g

using namespace c2g;

extern "C" {

const char* c2g_last_error(void) { return get_error(); }
const char* c2g_version(void) { return "c2g-mi355x 0.1 (gfx950)"; }

int64_t c2g_read_i32(const char* path, int32_t* buf, int64_t cap) { return read_raw(path, buf, cap); }
int64_t c2g_read_f32(const char* path, float* buf, int64_t cap) { return read_raw(path, buf, cap); }
int c2g_write_i32(const char* path, const int32_t* buf, int64_t n) { return write_raw(path, buf, n); }
int c2g_write_f32(const char* path, const float* buf, int64_t n) { return write_raw(path, buf, n); }

void c2g_free(void* p) { free(p); }

int c2g_write_dataset(const char* ds, int32_t V, const int32_t* vlist, const int32_t* elist,
                      const int32_t* wlist, const float* flist) {
    if (!ds || !vlist || V < 0) return fail(C2G_EINVAL, "c2g_write_dataset: bad arguments");
    std::string d = std::string(ds) + "/origin";
    int rc = mkdirs(d);
    if (rc) return rc;
    // The reference's other variant directories are created too, so that
    // purn-sssp / purn-ppr can write into them (src/purn-sssp.cpp:72-74).
    if ((rc = mkdirs(std::string(ds) + "/sssp"))) return rc;
    if ((rc = mkdirs(std::string(ds) + "/ppr"))) return rc;
    int64_t E = vlist[V];
    if ((rc = write_raw((d + "/vlist.bin").c_str(), vlist, (int64_t)V + 1))) return rc;
    if ((rc = write_raw((d + "/elist.bin").c_str(), elist, E))) return rc;
    if (wlist && (rc = write_raw((d + "/wlist.bin").c_str(), wlist, E))) return rc;
    if (flist && (rc = write_raw((d + "/flist.bin").c_str(), flist, E))) return rc;
    return C2G_OK;
}

}  // extern "C"
