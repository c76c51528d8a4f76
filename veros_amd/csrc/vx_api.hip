// Library-level entry points of libvxmix: version, limits, error messages.
#include "vx_common.hpp"

namespace vx {
std::string& last_error() {
    static thread_local std::string msg;
    return msg;
}
}  // namespace vx

extern "C" const char* vx_version(void) { return "vxmix 0.1 (gfx950)"; }

extern "C" int32_t vx_max_nz(void) { return vx::kMaxNz; }

extern "C" const char* vx_last_error(void) { return vx::last_error().c_str(); }
