// nh_server.hip -- continuous-batching server (placeholder entry points).
#include "../../include/necsim_hip.h"
#include "nh_internal.h"

extern "C" int nh_server_create(nh_ctx *, const nh_tables *, const nh_server_params *, nh_server **out) {
    if (out) *out = nullptr;
    return set_error(NH_E_STATE, "server not available");
}
extern "C" int nh_server_submit(nh_server *, const nh_request *, uint64_t *) { return set_error(NH_E_STATE, "server not available"); }
extern "C" int nh_server_tick(nh_server *, uint64_t *, uint64_t *) { return set_error(NH_E_STATE, "server not available"); }
extern "C" int nh_server_poll(nh_server *, nh_request_result *) { return set_error(NH_E_STATE, "server not available"); }
extern "C" int nh_server_request_events(nh_server *, uint64_t, nh_speciation_event *, size_t, size_t *) {
    return set_error(NH_E_STATE, "server not available");
}
extern "C" int nh_server_release(nh_server *, uint64_t) { return set_error(NH_E_STATE, "server not available"); }
extern "C" int nh_server_destroy(nh_server *) { return NH_OK; }
