// placeholder, filled in by the scatter implementation
#include "ebg_common.hpp"
extern "C" int ebg_scatter_release(ebg_ctx* ctx) { (void)ctx; return EBG_OK; }
