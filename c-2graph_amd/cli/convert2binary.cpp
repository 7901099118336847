// convert2binary <adjlist.txt> <dataset_path> (utils/convert2binary.cpp:86-135)
#include "cli_common.h"
int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "Usage: convert2binary <adjlist.txt> <dataset_path>\n"); return 2; }
    printf("Reading graph from %s\n", argv[1]);
    int rc = c2g_convert_adjlist(argv[1], argv[2]);
    return rc ? die(rc) : 0;
}
