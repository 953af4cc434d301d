// Serving loop (continuous batching) -- entry points.
#include <hip/hip_runtime.h>

#include "../../include/de_api.h"

extern "C" {

int de_server_create(size_t, int32_t, de_server** out) {
  if (out) *out = nullptr;
  return DE_ERR_UNSUPPORTED;
}
int de_server_submit(de_server*, const de_ode_batch*, uint64_t*) { return DE_ERR_UNSUPPORTED; }
int de_server_poll(de_server*, uint64_t, de_ode_result*, int32_t) { return DE_ERR_UNSUPPORTED; }
double de_server_latency(de_server*, uint64_t) { return -1.0; }
void de_server_destroy(de_server*) {}

}  // extern "C"
