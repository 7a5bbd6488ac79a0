// Command-line entry point: the reference's patternmatching_binary
// (src/patternmatching/patternmatching_binary.cc:13-42) running on the GPU.
#include "../../../include/prunejuice_host.h"

int main(int argc, char** argv) { return pjh_main(argc, argv); }
