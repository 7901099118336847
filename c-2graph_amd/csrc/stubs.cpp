// Entry points declared in c2g.h whose implementation has not landed yet.
#include "common.h"
using namespace c2g;
extern "C" {
int c2g_autotune(c2g_graph*, const int32_t*, int32_t, int32_t, c2g_tune*) { return fail(C2G_ENOTSUP, "not implemented"); }
}
