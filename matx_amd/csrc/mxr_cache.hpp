// Registration of per-module cache-release functions run by mxr_clear_caches,
// and the kernel-timing scope used around every launch.
#pragma once

#include <hip/hip_runtime.h>

#include "mxr.h"

namespace mxr {
void register_cache_clear(void (*fn)());

// Registers fn at static-initialisation time of the defining translation unit.
struct CacheClearRegistrar {
  explicit CacheClearRegistrar(void (*fn)()) { register_cache_clear(fn); }
};

// Device scratch of at least `bytes` for (current device, stream, tag);
// owned by the library, released by mxr_clear_caches.  A buffer that has to
// grow is replaced, and the old one is kept until mxr_clear_caches (a graph
// captured earlier may still use it).  Growing is refused while `s` is being
// captured (nullptr, with the reason in mxr_last_error): run the call once
// before capturing.  nullptr on failure.
void *stream_workspace(const char *tag, size_t bytes, hipStream_t s);
// Status for a null stream_workspace result: the capture refusal, or
// MXR_OUT_OF_MEMORY with a message naming fn.
mxr_status workspace_failure(const char *fn);

// Live-plan count: twiddle tables stay allocated through mxr_clear_caches
// while it is non-zero (plans hold raw pointers to them).
void plan_tables_retain();
void plan_tables_release();

// Brackets one kernel launch: profiling events when profiling is on, an
// INTERNAL trace range named after the kernel when the trace level asks for
// it, a TRACE log line when the log level does.
class ProfScope {
 public:
  ProfScope(const char *kernel, hipStream_t s);
  ~ProfScope();

 private:
  int slot_ = -1;
  bool traced_ = false;
  hipStream_t s_;
};
}  // namespace mxr
