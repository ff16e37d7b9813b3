// pfs_sim: the reference's simulation driver (src/main.cpp) on the GPU.  The
// driver lives in libpfs.so as pfs_sim_main so that it can also be called in
// process (tests, embedding).
#include "pfs.h"

int main(int argc, char **argv) { return pfs_sim_main(argc, argv); }
