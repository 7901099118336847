// ppr-base <dataset_path> <source> [--alpha A] [--eps E] [--reps N] [--out p.bin]
// (src/ppr/base/Main.cu:25-120)
#include "ppr_main.h"
int main(int argc, char** argv) { return ppr_main(argc, argv, 0, "ppr-base <dataset_path> <source> [--alpha A] [--eps E] [--out file]"); }
