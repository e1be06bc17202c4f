// Registration of per-module cache-release functions run by mxr_clear_caches,
// and the kernel-timing scope used around every launch.
#pragma once

#include <hip/hip_runtime.h>

namespace mxr {
void register_cache_clear(void (*fn)());

// Registers fn at static-initialisation time of the defining translation unit.
struct CacheClearRegistrar {
  explicit CacheClearRegistrar(void (*fn)()) { register_cache_clear(fn); }
};

// Device scratch of at least `bytes` for (current device, stream, tag);
// owned by the library, released by mxr_clear_caches.  nullptr on failure.
void *stream_workspace(const char *tag, size_t bytes, hipStream_t s);

// Brackets one kernel launch with profiling events when profiling is on.
class ProfScope {
 public:
  ProfScope(const char *kernel, hipStream_t s);
  ~ProfScope();

 private:
  int slot_ = -1;
  hipStream_t s_;
};
}  // namespace mxr
