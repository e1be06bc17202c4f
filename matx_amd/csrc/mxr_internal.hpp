// Internal helpers shared by the libmxr translation units: error reporting,
// tensor-descriptor checks, complex arithmetic and launch utilities.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "mxr.h"

namespace mxr {

// ----------------------------------------------------------------- errors
// The last error message is thread-local, like errno; the C ABI never throws.
void set_error(mxr_status st, const char *fmt, ...);
inline mxr_status ok() { return MXR_SUCCESS; }

#define MXR_REQUIRE(cond, status, ...)                     \
  do {                                                     \
    if (!(cond)) {                                         \
      ::mxr::set_error((status), __VA_ARGS__);             \
      return (status);                                     \
    }                                                      \
  } while (0)

#define MXR_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::mxr::set_error(MXR_HIP_ERROR, "%s failed: %s (%s:%d)", #expr,         \
                       hipGetErrorString(e_), __FILE__, __LINE__);           \
      return MXR_HIP_ERROR;                                                  \
    }                                                                        \
  } while (0)

// Wraps a C-ABI body so that no C++ exception crosses the boundary.
template <class F>
mxr_status guarded(const char *fn, F &&f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    set_error(MXR_OUT_OF_MEMORY, "%s: host allocation failed", fn);
    return MXR_OUT_OF_MEMORY;
  } catch (...) {
    set_error(MXR_NOT_SUPPORTED, "%s: unexpected exception", fn);
    return MXR_NOT_SUPPORTED;
  }
}

// ---------------------------------------------------------------- tensors
inline int64_t numel(const mxr_tensor &t) {
  int64_t n = 1;
  for (int d = 0; d < t.rank; ++d) n *= t.shape[d];
  return n;
}
inline size_t dtype_size(int dt) {
  switch (dt) {
    case MXR_C64: return 8;
    case MXR_F32: return 4;
    case MXR_I32: return 4;
    default: return 0;
  }
}
mxr_status check_tensor(const mxr_tensor *t, const char *name, int want_dtype,
                        int min_rank, int max_rank);
// [lo, hi) byte range touched by a view (for alias checks).
void byte_span(const mxr_tensor &t, uintptr_t *lo, uintptr_t *hi);
bool views_overlap(const mxr_tensor &a, const mxr_tensor &b);

// Batch-index decomposition carried into kernels: up to 7 batch dims.
struct BatchIdx {
  int32_t nd;
  int64_t shape[MXR_MAX_RANK];
  int64_t sa[MXR_MAX_RANK];  // strides of tensor A per batch dim
  int64_t sb[MXR_MAX_RANK];  // strides of tensor B per batch dim
};

__host__ __device__ inline void batch_offsets(const BatchIdx &bi, int64_t flat,
                                              int64_t *oa, int64_t *ob) {
  int64_t a = 0, b = 0;
  for (int d = bi.nd - 1; d >= 0; --d) {
    int64_t s = bi.shape[d];
    int64_t i = flat % s;
    flat /= s;
    a += i * bi.sa[d];
    b += i * bi.sb[d];
  }
  *oa = a;
  *ob = b;
}

// -------------------------------------------------------- device complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
// v if keep, else +0: a bit mask rather than a select, so that a load
// feeding it stays unconditional (the compiler turns a select between a
// load and a constant into a branch around the load, which serialises a
// thread's loads)
__device__ __forceinline__ float2 keep_if(float2 v, bool keep) {
  const int m = keep ? -1 : 0;
  return make_float2(__int_as_float(__float_as_int(v.x) & m), __int_as_float(__float_as_int(v.y) & m));
}

// Hamming window evaluated in float as MatX's generator does
// (include/matx/generators/hamming.h:96-99).
__device__ __forceinline__ float hamming_f32(int i, int n) {
  if (n <= 1) return 1.0f;
  return 0.54f - 0.46f * cosf(6.28318530717958647692f * (float)i / (float)(n - 1));
}

// ----------------------------------------------------------- launch util
// MI355X dispatches workgroups round-robin over its 8 XCDs.  Re-map the
// hardware id so that ids landing on one XCD get consecutive logical ids
// (neighbouring work shares that XCD's L2).  Bijective on [0, G).
__device__ __forceinline__ int64_t xcd_remap(int64_t g, int64_t G) {
  constexpr int64_t NX = 8;
  int64_t q = G / NX;
  if (g >= q * NX) return g;
  return (g % NX) * q + g / NX;
}

int device_id();
int cu_count();  // compute units of the current device (cached)
int max_lds_per_block();

}  // namespace mxr
