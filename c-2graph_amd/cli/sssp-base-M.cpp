// sssp-base-M <dataset_path> <source> <migration>: GPU for <migration> rounds,
// then the host Bellman-Ford lane (src/sssp/base-perk/Main.cu:29-260).
#include <iostream>
#include "sssp_main.h"
int main(int argc, char** argv) { return sssp_main(argc, argv, 2, "sssp-base-M <dataset_path> <source> <migration> [--reps N] [--out file]"); }
