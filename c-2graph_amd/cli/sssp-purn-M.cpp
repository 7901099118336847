// sssp-purn-M <dataset_path> <source> <migration> (src/sssp/purn-perk/Main.cu:22-270).
#include <iostream>
#include "sssp_main.h"
int main(int argc, char** argv) { return sssp_main(argc, argv, 3, "sssp-purn-M <dataset_path> <source> <migration> [--reps N] [--out file]"); }
