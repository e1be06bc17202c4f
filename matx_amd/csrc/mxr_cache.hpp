// Registration of per-module cache-release functions run by mxr_clear_caches.
#pragma once

namespace mxr {
void register_cache_clear(void (*fn)());

// Registers fn at static-initialisation time of the defining translation unit.
struct CacheClearRegistrar {
  explicit CacheClearRegistrar(void (*fn)()) { register_cache_clear(fn); }
};
}  // namespace mxr
