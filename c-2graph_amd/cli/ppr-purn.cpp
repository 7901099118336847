// ppr-purn <dataset_path> <source> (src/ppr/purn/Main.cu:26-140), on the purn-ppr graph.
#include "ppr_main.h"
int main(int argc, char** argv) { return ppr_main(argc, argv, 1, "ppr-purn <dataset_path> <source> [--alpha A] [--eps E] [--out file]"); }
