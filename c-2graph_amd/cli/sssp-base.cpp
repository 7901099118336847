// sssp-base <dataset_path> <source> [--reps N] [--out dist.bin] [--device D]
// (src/sssp/base/Main.cu:16-103)
#include <iostream>
#include "sssp_main.h"
int main(int argc, char** argv) { return sssp_main(argc, argv, 0, "sssp-base <dataset_path> <source> [--reps N] [--out file]"); }
