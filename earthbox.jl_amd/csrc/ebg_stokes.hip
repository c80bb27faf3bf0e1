// placeholder, filled in by the stokes implementation
#include "ebg_common.hpp"
extern "C" int ebg_stokes_release(ebg_ctx* ctx) { (void)ctx; return EBG_OK; }
