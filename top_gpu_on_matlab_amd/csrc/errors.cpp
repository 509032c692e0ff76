// Thread-local error state of the C ABI.
#include <cstring>

#include "tg_internal.h"

namespace tg {

static thread_local char g_id[128] = "";
static thread_local char g_msg[1024] = "";

void set_error(int, const char *id, const char *msg) {
    std::snprintf(g_id, sizeof(g_id), "%s", id ? id : "");
    std::snprintf(g_msg, sizeof(g_msg), "%s", msg ? msg : "");
}

void clear_error() {
    g_id[0] = 0;
    g_msg[0] = 0;
}

}  // namespace tg

extern "C" {
const char *tg_last_error_id(void) { return tg::g_id; }
const char *tg_last_error_message(void) { return tg::g_msg; }
const char *tg_version(void) { return "topgpu-mi355x 0.1 (gfx950)"; }
}
