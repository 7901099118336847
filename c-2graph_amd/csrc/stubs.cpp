// Entry points declared in c2g.h whose implementation has not landed yet.
#include "common.h"
using namespace c2g;
extern "C" {
int c2g_convert_adjlist(const char*, const char*) { return fail(C2G_ENOTSUP, "not implemented"); }
int c2g_autotune(c2g_graph*, const int32_t*, int32_t, int32_t, c2g_tune*) { return fail(C2G_ENOTSUP, "not implemented"); }
int c2g_server_start(c2g_graph*, const c2g_server_opts*, c2g_server**) { return fail(C2G_ENOTSUP, "not implemented"); }
int64_t c2g_server_submit(c2g_server*, int32_t, int32_t) { return fail(C2G_ENOTSUP, "not implemented"); }
int c2g_server_poll(c2g_server*, int64_t, int32_t, void*, c2g_result_info*) { return fail(C2G_ENOTSUP, "not implemented"); }
int c2g_server_stop(c2g_server*) { return fail(C2G_ENOTSUP, "not implemented"); }
}
