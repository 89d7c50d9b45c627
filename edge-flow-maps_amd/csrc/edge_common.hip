// Error reporting and library information for libedge_hip.so.
#include "edge_common.h"

namespace edge {

static thread_local char g_err[1024];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

}  // namespace edge

extern "C" {

// Thread-local message of the last failing call on this thread ("" if none).
const char* edge_last_error(void) { return edge::g_err; }

// ABI version of this library (bumped when an exported signature changes).
int edge_abi_version(void) { return 1; }

// Synchronise a stream (the reference's ti.sync(), util/timer.py:19,27).
int edge_stream_sync(void* stream) {
  edge::clear_error();
  EDGE_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return edge::EDGE_OK;
}

}  // extern "C"
