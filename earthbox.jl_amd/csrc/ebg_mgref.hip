// placeholder, filled in by the mgref implementation
#include "ebg_common.hpp"
extern "C" int ebg_mgref_release(ebg_ctx* ctx) { (void)ctx; return EBG_OK; }
