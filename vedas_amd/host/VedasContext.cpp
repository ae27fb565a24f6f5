#include "VedasContext.h"

static thread_local VedasContext *g_current = nullptr;

VedasContext::VedasContext(int device) : device_(device) {
    int rc = vd_ctx_create(device, &ctx_);
    if (rc != VD_OK)
        throw VedasError(rc, "vd_ctx_create failed for device " + std::to_string(device));
}

VedasContext::~VedasContext() {
    if (g_current == this) g_current = nullptr;
    if (ctx_) vd_ctx_destroy(ctx_);
}

void VedasContext::check(int rc, const char *what) const {
    if (rc != VD_OK)
        throw VedasError(rc, std::string(what) + ": " + vd_ctx_last_error(ctx_));
}

void VedasContext::sync() const { check(vd_ctx_sync(ctx_), "vd_ctx_sync"); }

VedasContext *VedasContext::current() { return g_current; }
void VedasContext::setCurrent(VedasContext *c) { g_current = c; }

void VedasContext::resetStats() {
    upload_ns = join_ns = swap_index_ns = eliminate_duplicate_ns = eif_ns = 0;
    eif_count = 0;
    join_rows = select_rows = swap_rows = 0;
    driven_joins = 0;
    filtered_rows = 0;
    exe_log.clear();
}
