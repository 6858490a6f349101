// ckt_server.cpp -- continuous batching of circuit simulations.
#include "ckt_engine.h"

using namespace ckt;

struct ckt_server {
  Engine eng;
};

extern "C" {

int ckt_server_create(int device, const ckt_server_config_t* config, ckt_server_t** out) {
  (void)device;
  (void)config;
  if (out) *out = nullptr;
  return fail(CKT_ERR_UNSUPPORTED, "server not built yet");
}
int ckt_server_destroy(ckt_server_t* s) {
  delete s;
  return CKT_OK;
}
int ckt_server_run_trace(ckt_server_t*, const ckt_request_t*, uint32_t, ckt_request_result_t*,
                         ckt_server_report_t*) {
  return fail(CKT_ERR_UNSUPPORTED, "server not built yet");
}
int ckt_server_set_keep_state(ckt_server_t*, int) { return fail(CKT_ERR_UNSUPPORTED, "server not built yet"); }
int ckt_server_fetch(ckt_server_t*, uint32_t, ckt_node_t*, ckt_wire_t*) {
  return fail(CKT_ERR_UNSUPPORTED, "server not built yet");
}

}  // extern "C"
