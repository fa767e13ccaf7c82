// server.hip -- request loop (continuous batching).  Not implemented yet: every entry point
// reports DTX_ERR_UNSUPPORTED.
#include "dtx_internal.h"

struct dtx_server {
    int unused;
};

extern "C" {
int dtx_server_create(const dtx_server_config*, dtx_server** out) {
    if (out) *out = nullptr;
    return DTX_ERR_UNSUPPORTED;
}
int dtx_server_submit(dtx_server*, const dtx_request*, int64_t*) { return DTX_ERR_UNSUPPORTED; }
int dtx_server_poll(dtx_server*, int64_t, dtx_request_result*) { return DTX_ERR_UNSUPPORTED; }
int dtx_server_fetch(dtx_server*, int64_t, double*, int64_t) { return DTX_ERR_UNSUPPORTED; }
int dtx_server_drain(dtx_server*) { return DTX_ERR_UNSUPPORTED; }
void dtx_server_destroy(dtx_server*) {}
const char* dtx_server_last_error(const dtx_server*) { return "serving loop not built"; }
}
