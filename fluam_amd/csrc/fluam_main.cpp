// fluam_main.cpp -- command-line driver with fluam's interface: `fluam [data.main]`
// (README.md:49 of the reference; dispatch of src/main.cpp:28-112).
#include "../../include/fhd.h"

int main(int argc, char** argv) { return fhd_main(argc > 1 ? argv[1] : "data.main") ? 0 : 1; }
