// tbh_serve.cpp -- continuous-batching request scheduler (tbh_submit/poll).
#include "tbh_runtime.hpp"

namespace tbh {
struct ServeState {};
}  // namespace tbh

extern "C" {
int tbh_serve_start(tbh_ctx* ctx, size_t) {
  (void)ctx;
  tbh::set_error("serving not built yet");
  return TBH_EINVAL;
}
int tbh_serve_stop(tbh_ctx* ctx) {
  (void)ctx;
  return 0;
}
int tbh_submit(tbh_ctx*, const tbh_request*, uint64_t*) {
  tbh::set_error("serving not started");
  return TBH_EINVAL;
}
int tbh_poll(tbh_ctx*, uint64_t, void*, tbh_status*, double*, double*) {
  tbh::set_error("serving not started");
  return TBH_EINVAL;
}
int tbh_wait(tbh_ctx*, uint64_t, void*, tbh_status*) {
  tbh::set_error("serving not started");
  return TBH_EINVAL;
}
int tbh_cancel(tbh_ctx*, uint64_t) {
  tbh::set_error("serving not started");
  return TBH_EINVAL;
}
}
