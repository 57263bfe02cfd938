// Library-provided main (src/charming.cu:50-263): charm::main_host, then
// chm_init, chm_run, chm_finalize.  Linked into chare-program executables.
#include "charming_amd.h"

int main(int argc, char** argv) { return chm_main(argc, argv); }
