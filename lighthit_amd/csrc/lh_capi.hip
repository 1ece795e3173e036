// Error reporting and version of the lighthit_amd C-ABI.
#include "lh_common.h"

#include <cstring>

namespace lh {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
}

}  // namespace lh

extern "C" {

const char* lh_last_error(void) { return lh::g_error; }

void lh_clear_error(void) { lh::g_error[0] = '\0'; }

int lh_abi_version(void) { return LH_ABI_VERSION; }

}
