// ppr-purn-M <dataset_path> <source> <iter> (src/ppr/w-purn-perk/Main.cu).
#include "ppr_main.h"
int main(int argc, char** argv) { return ppr_main(argc, argv, 3, "ppr-purn-M <dataset_path> <source> <iter> [--handoff N] [--out file]"); }
