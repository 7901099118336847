// sssp-purn <dataset_path> <source> [--reps N] [--out dist.bin]
// (src/sssp/purn/Main.cu:18-126); distances equal those of the full graph.
#include <iostream>
#include "sssp_main.h"
int main(int argc, char** argv) { return sssp_main(argc, argv, 1, "sssp-purn <dataset_path> <source> [--reps N] [--out file]"); }
