// libmxr runtime: thread-local errors, descriptor checks, per-device caches
// of twiddle tables (the analogue of MatX's plan cache,
// include/matx/core/cache.h:167-291), and mxr_clear_caches.
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <mutex>
#include <vector>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "mxr_trace.hpp"

namespace mxr {

static thread_local std::string g_last_error;

void set_error(mxr_status st, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = std::string("[mxr status ") + std::to_string((int)st) + "] " + buf;
  log_msg(LOG_ERROR, "%s", g_last_error.c_str());
}

mxr_status check_tensor(const mxr_tensor *t, const char *name, int want_dtype, int min_rank,
                        int max_rank) {
  MXR_REQUIRE(t != nullptr, MXR_INVALID_PARAMETER, "%s: null descriptor", name);
  MXR_REQUIRE(t->rank >= min_rank && t->rank <= max_rank, MXR_INVALID_DIM,
              "%s: rank %d outside [%d, %d]", name, t->rank, min_rank, max_rank);
  if (want_dtype) {
    MXR_REQUIRE(t->dtype == want_dtype, MXR_INVALID_TYPE, "%s: dtype %d, expected %d", name,
                t->dtype, want_dtype);
  }
  MXR_REQUIRE(dtype_size(t->dtype) != 0, MXR_INVALID_TYPE, "%s: unknown dtype %d", name, t->dtype);
  for (int d = 0; d < t->rank; ++d) {
    MXR_REQUIRE(t->shape[d] >= 0, MXR_INVALID_SIZE, "%s: negative extent in dim %d", name, d);
  }
  MXR_REQUIRE(t->data != nullptr || numel(*t) == 0, MXR_INVALID_PARAMETER, "%s: null data", name);
  return MXR_SUCCESS;
}

void byte_span(const mxr_tensor &t, uintptr_t *lo, uintptr_t *hi) {
  int64_t mn = 0, mx = 0;
  for (int d = 0; d < t.rank; ++d) {
    if (t.shape[d] == 0) {
      *lo = *hi = (uintptr_t)t.data;
      return;
    }
    int64_t e = (t.shape[d] - 1) * t.stride[d];
    if (e < 0) mn += e;
    else mx += e;
  }
  const size_t es = dtype_size(t.dtype);
  *lo = (uintptr_t)t.data + (intptr_t)(mn * (int64_t)es);
  *hi = (uintptr_t)t.data + (uintptr_t)((mx + 1) * (int64_t)es);
}

bool views_overlap(const mxr_tensor &a, const mxr_tensor &b) {
  uintptr_t alo, ahi, blo, bhi;
  byte_span(a, &alo, &ahi);
  byte_span(b, &blo, &bhi);
  return alo < bhi && blo < ahi;
}

int device_id() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

int cu_count() {
  static std::mutex mu;
  static std::map<int, int> cache;
  std::lock_guard<std::mutex> g(mu);
  int d = device_id();
  auto it = cache.find(d);
  if (it != cache.end()) return it->second;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || v < 1) v = 256;
  cache[d] = v;
  return v;
}

int max_lds_per_block() {
  static std::mutex mu;
  static std::map<int, int> cache;
  std::lock_guard<std::mutex> g(mu);
  int d = device_id();
  auto it = cache.find(d);
  if (it != cache.end()) return it->second;
  int v = 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, d) != hipSuccess) v = 65536;
  cache[d] = v;
  return v;
}

void build_twiddles(int N, RadixSeq s, float2 *out) {
  const double two_pi = 6.283185307179586476925286766559;
  int64_t Ls = 1;
  for (int p = 0; p < s.n; ++p) {
    const int R = s.r[p];
    if (Ls > 1) {
      float2 *t = out + tw_offset(s, p);
      const int64_t len = Ls * R;
      for (int64_t k = 0; k < Ls; ++k) {
        for (int m = 0; m < R; ++m) {
          // reduce m*k modulo len before scaling: the angle index stays exact
          const int64_t e = (m * k) % len;
          const double a = -two_pi * (double)e / (double)len;
          t[k * R + m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
      }
    }
    Ls *= R;
  }
  (void)N;
}

// --------------------------------------------------------- device caches
namespace {
struct TwKey {
  int dev, n, rev;
  bool operator<(const TwKey &o) const {
    return dev != o.dev ? dev < o.dev : (n != o.n ? n < o.n : rev < o.rev);
  }
};
std::mutex g_tw_mu;
std::map<TwKey, float2 *> g_tw;

std::mutex g_hooks_mu;
std::vector<void (*)()> &clear_hooks() {
  static std::vector<void (*)()> h;
  return h;
}
}  // namespace

void register_cache_clear(void (*fn)()) {
  std::lock_guard<std::mutex> g(g_hooks_mu);
  clear_hooks().push_back(fn);
}

const float2 *twiddles_device(int N, bool rev) {
  const int dev = device_id();
  std::lock_guard<std::mutex> g(g_tw_mu);
  TwKey key{dev, N, rev ? 1 : 0};
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  RadixSeq s = radix_seq(N);
  if (s.n < 0) return nullptr;
  if (rev) s = reversed(s);
  const int64_t sz = tw_size(s) < 2 ? 2 : tw_size(s);
  std::vector<float2> host((size_t)sz, make_float2(1.f, 0.f));
  build_twiddles(N, s, host.data());
  float2 *d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * sz) != hipSuccess) return nullptr;
  // a synchronous copy: tables are built once per (device, size), before use
  if (hipMemcpy(d, host.data(), sizeof(float2) * sz, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_tw[key] = d;
  log_msg(LOG_DEBUG, "twiddle table: n=%d%s, %lld entries on device %d", N, rev ? " (reverse order)" : "",
          (long long)sz, dev);
  return d;
}

// ------------------------------------------------------------ profiling
namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
  bool used = false;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof;  // pool; records reused after reset
size_t g_prof_n = 0;
bool g_prof_on = false;
}  // namespace

ProfScope::ProfScope(const char *kernel, hipStream_t s) : s_(s) {
  traced_ = trace_push(kernel, MXR_TRACE_INTERNAL);
  if (log_on(LOG_TRACE)) log_msg(LOG_TRACE, "launch %s on stream %p", kernel, (void *)s);
  std::lock_guard<std::mutex> g(g_prof_mu);
  if (!g_prof_on) return;
  if (g_prof_n == g_prof.size()) {
    ProfRec r;
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
    g_prof.push_back(r);
  }
  ProfRec &r = g_prof[g_prof_n];
  r.name = kernel;
  r.used = true;
  if (hipEventRecord(r.a, s) != hipSuccess) return;
  slot_ = (int)g_prof_n++;
}

ProfScope::~ProfScope() {
  if (slot_ >= 0) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    (void)hipEventRecord(g_prof[slot_].b, s_);
  }
  if (traced_) trace_pop();
}

namespace {
std::map<TwKey, float4 *> g_pb;
}

const float4 *pow_base_device(int N, int Ls) {
  const int dev = device_id();
  std::lock_guard<std::mutex> g(g_tw_mu);
  TwKey key{dev, N, Ls};
  auto it = g_pb.find(key);
  if (it != g_pb.end()) return it->second;
  const double two_pi = 6.283185307179586476925286766559;
  const int64_t len = (int64_t)16 * Ls;
  std::vector<float4> host((size_t)Ls * 2);
  for (int64_t k = 0; k < Ls; ++k) {
    float2 w[4];
    const int pw[4] = {1, 2, 4, 8};
    for (int i = 0; i < 4; ++i) {
      const int64_t e = (pw[i] * k) % len;
      const double ang = -two_pi * (double)e / (double)len;
      w[i] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    host[2 * k] = make_float4(w[0].x, w[0].y, w[1].x, w[1].y);
    host[2 * k + 1] = make_float4(w[2].x, w[2].y, w[3].x, w[3].y);
  }
  float4 *d = nullptr;
  if (hipMalloc(&d, sizeof(float4) * host.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, host.data(), sizeof(float4) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_pb[key] = d;
  return d;
}

// Per (device, stream, tag) scratch buffers that only grow: partial sums of
// reductions that a launch on that stream writes and a later launch on the
// same stream consumes.  A buffer that has to grow is replaced and the old
// one retired until mxr_clear_caches (see stream_workspace).
namespace {
std::mutex g_ws_mu;
struct WsBuf {
  void *p = nullptr;
  size_t bytes = 0;
};
std::map<std::tuple<int, uintptr_t, std::string>, WsBuf> g_ws;

// Superseded buffers (device, pointer) stay allocated until mxr_clear_caches:
// a hipGraph captured earlier may still write them on replay.
std::vector<std::pair<int, void *>> g_ws_retired;
thread_local bool t_ws_capture_refused = false;
}  // namespace

mxr_status workspace_failure(const char *fn) {
  if (t_ws_capture_refused) {
    t_ws_capture_refused = false;
    return MXR_NOT_SUPPORTED;  // message already set by stream_workspace
  }
  set_error(MXR_OUT_OF_MEMORY, "%s: workspace allocation failed", fn);
  return MXR_OUT_OF_MEMORY;
}

void *stream_workspace(const char *tag, size_t bytes, hipStream_t s) {
  const int dev = device_id();
  const auto key = std::make_tuple(dev, (uintptr_t)s, std::string(tag));
  std::lock_guard<std::mutex> g(g_ws_mu);
  WsBuf &b = g_ws[key];
  t_ws_capture_refused = false;
  if (b.bytes >= bytes && b.p) return b.p;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
    set_error(MXR_NOT_SUPPORTED,
              "workspace '%s' must grow to %zu bytes while the stream is being captured: run "
              "the call once outside the capture first", tag, bytes);
    t_ws_capture_refused = true;
    return nullptr;
  }
  if (b.p) g_ws_retired.emplace_back(dev, b.p);
  b.p = nullptr;
  b.bytes = 0;
  log_msg(LOG_DEBUG, "workspace '%s' on stream %p: %zu bytes", tag, (void *)s, bytes);
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    return nullptr;
  }
  b.bytes = bytes;
  return b.p;
}

static void clear_workspaces() {
  std::lock_guard<std::mutex> g(g_ws_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_ws) {
    (void)hipSetDevice(std::get<0>(kv.first));
    (void)hipFree(kv.second.p);
  }
  for (auto &r : g_ws_retired) {
    (void)hipSetDevice(r.first);
    (void)hipFree(r.second);
  }
  (void)hipSetDevice(cur);
  g_ws.clear();
  g_ws_retired.clear();
}
CacheClearRegistrar g_reg_ws(clear_workspaces);

// Plans (pulse-compression and ragged) keep raw pointers to the twiddle and
// power-base tables; while any plan is alive mxr_clear_caches leaves the
// tables in place (a few hundred KiB), so a live plan never reads freed memory.
namespace {
int g_live_plans = 0;  // guarded by g_tw_mu
}
void plan_tables_retain() {
  std::lock_guard<std::mutex> g(g_tw_mu);
  ++g_live_plans;
}
void plan_tables_release() {
  std::lock_guard<std::mutex> g(g_tw_mu);
  if (g_live_plans > 0) --g_live_plans;
}

static void clear_twiddles() {
  std::lock_guard<std::mutex> g(g_tw_mu);
  if (g_live_plans > 0) return;
  for (auto &kv : g_tw) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first.dev);
    (void)hipFree(kv.second);
    (void)hipSetDevice(cur);
  }
  g_tw.clear();
  for (auto &kv : g_pb) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first.dev);
    (void)hipFree(kv.second);
    (void)hipSetDevice(cur);
  }
  g_pb.clear();
}

}  // namespace mxr

extern "C" {

const char *mxr_last_error(void) { return mxr::g_last_error.c_str(); }

int mxr_version(void) { return 100; }

mxr_status mxr_profile_enable(int on) {
  std::lock_guard<std::mutex> g(mxr::g_prof_mu);
  mxr::g_prof_on = on != 0;
  return MXR_SUCCESS;
}

mxr_status mxr_profile_reset(void) {
  std::lock_guard<std::mutex> g(mxr::g_prof_mu);
  for (size_t i = 0; i < mxr::g_prof_n; ++i) mxr::g_prof[i].used = false;
  mxr::g_prof_n = 0;
  return MXR_SUCCESS;
}

mxr_status mxr_profile_query(const char *name, int64_t *count, double *total_ms) {
  return mxr::guarded("mxr_profile_query", [&]() -> mxr_status {
    MXR_REQUIRE(name && count && total_ms, MXR_INVALID_PARAMETER, "profile_query: null argument");
    std::lock_guard<std::mutex> g(mxr::g_prof_mu);
    int64_t n = 0;
    double t = 0.0;
    for (size_t i = 0; i < mxr::g_prof_n; ++i) {
      auto &r = mxr::g_prof[i];
      if (r.name.find(name) == std::string::npos) continue;
      MXR_HIP_TRY(hipEventSynchronize(r.b));
      float ms = 0.f;
      MXR_HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
      t += ms;
      ++n;
    }
    *count = n;
    *total_ms = t;
    return MXR_SUCCESS;
  });
}

mxr_status mxr_clear_caches(void) {
  return mxr::guarded("mxr_clear_caches", [&]() -> mxr_status {
    MXR_HIP_TRY(hipDeviceSynchronize());
    std::vector<void (*)()> hooks;
    {
      std::lock_guard<std::mutex> g(mxr::g_hooks_mu);
      hooks = mxr::clear_hooks();
    }
    for (auto fn : hooks) fn();
    mxr::clear_twiddles();
    mxr::log_msg(mxr::LOG_DEBUG, "caches cleared");
    return MXR_SUCCESS;
  });
}

mxr_status mxr_tensor_from_dlpack(mxr_tensor *out, const mxr_dl_tensor *src) {
  return mxr::guarded("mxr_tensor_from_dlpack", [&]() -> mxr_status {
    MXR_REQUIRE(out != nullptr && src != nullptr, MXR_INVALID_PARAMETER,
                "tensor_from_dlpack: null argument");
    constexpr int32_t kDLROCM = 10;
    MXR_REQUIRE(src->device_type == kDLROCM, MXR_NOT_SUPPORTED,
                "tensor_from_dlpack: device type %d is not ROCm device memory (10)",
                (int)src->device_type);
    MXR_REQUIRE(src->ndim >= 0 && src->ndim <= MXR_MAX_RANK, MXR_INVALID_DIM,
                "tensor_from_dlpack: rank %d not in [0, %d]", (int)src->ndim, MXR_MAX_RANK);
    MXR_REQUIRE(src->dtype_lanes == 1, MXR_INVALID_TYPE, "tensor_from_dlpack: %d lanes",
                (int)src->dtype_lanes);
    int dt = 0;
    if (src->dtype_code == 5 && src->dtype_bits == 64) dt = MXR_C64;
    else if (src->dtype_code == 2 && src->dtype_bits == 32) dt = MXR_F32;
    else if (src->dtype_code == 0 && src->dtype_bits == 32) dt = MXR_I32;
    MXR_REQUIRE(dt != 0, MXR_INVALID_TYPE,
                "tensor_from_dlpack: dtype (code %d, %d bits) is not complex64, float32 or int32",
                (int)src->dtype_code, (int)src->dtype_bits);
    MXR_REQUIRE(src->ndim == 0 || src->shape != nullptr, MXR_INVALID_PARAMETER,
                "tensor_from_dlpack: null shape");
    const size_t esz = dt == MXR_C64 ? 8 : 4;
    MXR_REQUIRE(src->byte_offset % esz == 0, MXR_INVALID_PARAMETER,
                "tensor_from_dlpack: byte offset %llu is not a multiple of the element size",
                (unsigned long long)src->byte_offset);
    mxr_tensor t{};
    t.data = static_cast<char *>(src->data) + src->byte_offset;
    t.dtype = dt;
    t.rank = src->ndim;
    int64_t compact = 1;
    for (int d = src->ndim - 1; d >= 0; --d) {
      MXR_REQUIRE(src->shape[d] >= 0, MXR_INVALID_SIZE, "tensor_from_dlpack: negative extent");
      t.shape[d] = src->shape[d];
      t.stride[d] = src->strides ? src->strides[d] : compact;
      compact *= src->shape[d];
    }
    *out = t;
    return MXR_SUCCESS;
  });
}

namespace {
// One exported view: the DLManagedTensor first, then its arrays and the
// producer's release callback.
struct DlExport {
  mxr_dl_managed m;
  int64_t shape[MXR_MAX_RANK];
  int64_t strides[MXR_MAX_RANK];
  void (*release)(void *);
  void *release_ctx;
};
void dl_export_deleter(mxr_dl_managed *self) {
  if (!self) return;
  DlExport *e = reinterpret_cast<DlExport *>(self);
  void (*rel)(void *) = e->release;
  void *ctx = e->release_ctx;
  delete e;
  if (rel) rel(ctx);
}
}  // namespace

mxr_status mxr_tensor_to_dlpack(const mxr_tensor *t, int32_t device_type, int32_t device_id,
                                void (*release)(void *), void *release_ctx, mxr_dl_managed **out) {
  return mxr::guarded("mxr_tensor_to_dlpack", [&]() -> mxr_status {
    MXR_REQUIRE(t != nullptr && out != nullptr, MXR_INVALID_PARAMETER,
                "tensor_to_dlpack: null argument");
    MXR_REQUIRE(device_type == 10 || device_type == 11, MXR_NOT_SUPPORTED,
                "tensor_to_dlpack: device type %d is not ROCm memory (10 device, 11 pinned host)",
                (int)device_type);
    mxr_status st = mxr::check_tensor(t, "tensor_to_dlpack", 0, 0, MXR_MAX_RANK);
    if (st) return st;
    DlExport *e = new DlExport{};
    mxr_dl_tensor &d = e->m.dl_tensor;
    d.data = t->data;
    d.device_type = device_type;
    d.device_id = device_id;
    d.ndim = t->rank;
    d.dtype_lanes = 1;
    if (t->dtype == MXR_C64) {
      d.dtype_code = 5;
      d.dtype_bits = 64;
    } else if (t->dtype == MXR_F32) {
      d.dtype_code = 2;
      d.dtype_bits = 32;
    } else {
      d.dtype_code = 0;
      d.dtype_bits = 32;
    }
    for (int i = 0; i < t->rank; ++i) {
      e->shape[i] = t->shape[i];
      e->strides[i] = t->stride[i];
    }
    d.shape = e->shape;
    d.strides = e->strides;
    d.byte_offset = 0;
    e->m.manager_ctx = e;
    e->m.deleter = dl_export_deleter;
    e->release = release;
    e->release_ctx = release_ctx;
    *out = &e->m;
    return MXR_SUCCESS;
  });
}

}  // extern "C"
