// cubble-cli: the one-line main of the drop-in command line (reference
// src/main.cpp:19-25).
#include "cubble_amd.h"

int main(int argc, char **argv) { return cubble_run(argc, argv); }
