// version.hip -- library identification.
#include "mvx_common.hpp"

extern "C" const char *mvx_version(void) { return "mivisionx_amd 0.1 (gfx950)"; }
