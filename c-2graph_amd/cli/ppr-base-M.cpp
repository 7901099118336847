// ppr-base-M <dataset_path> <source> <iter> (src/ppr/w-base-perk/Main.cu:29-240).
#include "ppr_main.h"
int main(int argc, char** argv) { return ppr_main(argc, argv, 2, "ppr-base-M <dataset_path> <source> <iter> [--handoff N] [--out file]"); }
