// Shared helpers of the command-line tools.
#pragma once
#include <c2g.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

inline int die(int rc) {
    fprintf(stderr, "error %d: %s\n", rc, c2g_last_error());
    return 1;
}
