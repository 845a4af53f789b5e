// Batched fills and graph capture (stubs until implemented).
#include "api_common.hpp"
#include "plan.hpp"

using namespace tah;

extern "C" {

tah_status tah_fill_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, void* stream) {
  (void)ctx; (void)n; (void)plans; (void)works; (void)factors; (void)stream;
  set_last_error("tah_fill_batched is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_create(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, tah_graph** out) {
  (void)ctx; (void)n; (void)plans; (void)works; (void)factors; (void)out;
  set_last_error("tah_graph_create is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_launch(tah_graph* g, void* stream) {
  (void)g; (void)stream;
  set_last_error("tah_graph_launch is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_destroy(tah_graph* g) {
  (void)g;
  return TAH_OK;
}

tah_status tah_fill_adjoint(tah_plan* p, double* adj, const double* factors, void* stream) {
  (void)p; (void)adj; (void)factors; (void)stream;
  set_last_error("tah_fill_adjoint is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_tune(tah_plan* p, double* work, const double* factors, const char* cache_path) {
  (void)p; (void)work; (void)factors; (void)cache_path;
  set_last_error("tah_tune is not implemented yet");
  return TAH_EUNSUPPORTED;
}

}  // extern "C"
