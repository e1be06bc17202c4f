// matx_amd.h -- C++ front end with MatX's operator API over libmxr.so.
//
// Header-only C++17/20: a MatX-style program (tensor_t, make_tensor,
// slice / Permute / Clone views, (lhs = op).run(exec), fft / ifft / fft_conv /
// conv1d / conv2d / sum, elementwise operators and generators) compiles with
// g++ (host only) or hipcc against this header and links libmxr.so; every
// built-in operation reaches the gfx950 kernels through the C ABI in mxr.h.
// The executor is HIPExecutor (a hipStream_t), the analogue of cudaExecutor
// (include/matx/executors/cuda.h:57-231 in MatX).
//
// User-defined operators derived from BaseOp<Op> (MatX's custom-operator
// form, e.g. calcDets in examples/simple_radar_pipeline.h:50-150) carry their
// own __device__ operator(); their run() launches a kernel instantiated in
// the user's translation unit, so that unit is compiled with hipcc.
//
// Differences from MatX, by design:
//  * elementwise expressions are lowered at run() time to one postfix program
//    evaluated by one kernel (mxr_ew_eval) instead of a compile-time template;
//  * a transform whose input is an expression materialises it into a
//    temporary first (MatX allocates the same temporaries,
//    include/matx/operators/fft.h:493-505);
//  * errors throw matxException carrying the mxr_status and message.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MATX_AMD_HD __host__ __device__
#else
#include <hip/hip_runtime_api.h>
#define MATX_AMD_HD
#endif

#include <cmath>

#include <array>
#include <complex>
#include <cstdint>
#include <cctype>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mxr.h"

// MatX user code spells the complex type, fixed-size arrays and scalar math
// with libcu++'s names (cuda::std::complex<float>, cuda::std::array,
// cuda::std::pow in examples/simple_radar_pipeline.h:94,118,156).  On this
// toolchain they are the C++ standard library's, whose math functions hipcc
// also provides in device code.
namespace cuda {
namespace std {
using ::std::array;
template <typename T>
using complex = ::std::complex<T>;
using ::std::abs;
using ::std::exp;
using ::std::log;
using ::std::pow;
using ::std::sqrt;
}  // namespace std
}  // namespace cuda

// Macros MatX operator code uses (include/matx/core/defines.h).
#ifndef __MATX_INLINE__
#define __MATX_INLINE__ inline
#endif
#ifndef __MATX_HOST__
#define __MATX_HOST__ __host__
#endif
#ifndef __MATX_DEVICE__
#define __MATX_DEVICE__ __device__
#endif

namespace matx {

using index_t = int64_t;
using cfloat = std::complex<float>;  // interleaved re,im: the mxr C64 layout

constexpr index_t matxKeepDim = std::numeric_limits<index_t>::max();
constexpr index_t matxEnd = std::numeric_limits<index_t>::max() - 1;
constexpr index_t matxDropDim = std::numeric_limits<index_t>::max() - 2;

enum matxConvCorrMode_t { MATX_C_MODE_FULL = 0, MATX_C_MODE_SAME = 1, MATX_C_MODE_VALID = 2 };
enum matxConvCorrMethod_t { MATX_C_METHOD_DIRECT = 0, MATX_C_METHOD_FFT = 1 };
enum class FFTNorm { BACKWARD = 0, FORWARD = 1, ORTHO = 2 };
enum matxMemorySpace_t { MATX_MANAGED_MEMORY, MATX_DEVICE_MEMORY, MATX_HOST_MALLOC_MEMORY };

class matxException : public std::runtime_error {
 public:
  matxException(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(mxr_status s) {
  if (s != MXR_SUCCESS) throw matxException((int)s, mxr_last_error());
}
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw matxException(MXR_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}
// Whether managed memory lives where the current device computes on it
// (pages migrate on demand); otherwise hipMallocManaged falls back to host
// memory.
inline bool managed_memory_migrates() {
  int dev = 0, v = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeConcurrentManagedAccess, dev) != hipSuccess) return false;
  return v != 0;
}
template <typename T>
struct dtype_of;
template <>
struct dtype_of<cfloat> {
  static constexpr int value = MXR_C64;
};
template <>
struct dtype_of<float> {
  static constexpr int value = MXR_F32;
};
template <>
struct dtype_of<int> {
  static constexpr int value = MXR_I32;
};
}  // namespace detail

// ------------------------------------------------------- trace ranges, log
// MatX's NVTX levels and macros (include/matx/core/nvtx.h:48-168) over the
// library's roctx ranges (mxr_trace_*), and its MATX_LOG_* macros
// (include/matx/core/log.h:389-401) over the library log (mxr_log_*).  The
// log macros take MatX's "{}" placeholders, filled in order by operator<<.
enum matx_nvxtLogLevels {
  MATX_NVTX_LOG_NONE = 0,
  MATX_NVTX_LOG_USER = 1,
  MATX_NVTX_LOG_API = 2,
  MATX_NVTX_LOG_INTERNAL = 3,
  MATX_NVTX_LOG_ALL = 0xFFFFFF
};

namespace detail {
// A nested range that ends with the scope (MATX_NVTX_START).
class TraceScope {
 public:
  TraceScope(const char *function, const std::string &message = "", int level = MATX_NVTX_LOG_USER)
      : opened_(mxr_trace_push(message.empty() ? function : message.c_str(), level) != 0) {}
  TraceScope(const TraceScope &) = delete;
  TraceScope &operator=(const TraceScope &) = delete;
  ~TraceScope() {
    if (opened_) mxr_trace_pop();
  }

 private:
  bool opened_;
};

// A range ended by id (MATX_NVTX_START_RANGE); id < 0 picks a fresh one.
inline int trace_range_start(const char *function, const std::string &message = "",
                             int level = MATX_NVTX_LOG_USER, int id = -1) {
  int i = id;
  check(mxr_trace_range_start(message.empty() ? function : message.c_str(), level, &i));
  return i;
}

inline void log_fill(std::ostringstream &o, const char *f) { o << f; }
template <typename A, typename... R>
void log_fill(std::ostringstream &o, const char *f, const A &a, const R &...r) {
  const char *p = std::strstr(f, "{}");
  if (!p) {
    o << f;
    return;
  }
  o.write(f, p - f);
  o << a;
  log_fill(o, p + 2, r...);
}
template <typename... A>
void log_at(int level, const char *fmt, const A &...a) {
  const int cur = mxr_log_get_level();
  if (cur >= 6 || level < cur) return;
  std::ostringstream o;
  log_fill(o, fmt, a...);
  mxr_log_write(level, o.str().c_str());
}
}  // namespace detail
}  // namespace matx

#define MATX_AMD_CAT2(a, b) a##b
#define MATX_AMD_CAT(a, b) MATX_AMD_CAT2(a, b)
#define MATX_NVTX_START(...) ::matx::detail::TraceScope MATX_AMD_CAT(matx_trace_, __COUNTER__)(__FUNCTION__, __VA_ARGS__);
#define MATX_NVTX_START_RANGE(...) ::matx::detail::trace_range_start(__FUNCTION__, __VA_ARGS__);
#define MATX_NVTX_END_RANGE(id) ::matx::detail::check(mxr_trace_range_end(id));
#define MATX_NVTX_SET_LOG_LEVEL(level) ::matx::detail::check(mxr_trace_set_level((int)(level)));
#define MATX_LOG_TRACE(...) ::matx::detail::log_at(0, __VA_ARGS__)
#define MATX_LOG_DEBUG(...) ::matx::detail::log_at(1, __VA_ARGS__)
#define MATX_LOG_INFO(...) ::matx::detail::log_at(2, __VA_ARGS__)
#define MATX_LOG_WARN(...) ::matx::detail::log_at(3, __VA_ARGS__)
#define MATX_LOG_ERROR(...) ::matx::detail::log_at(4, __VA_ARGS__)
#define MATX_LOG_FATAL(...) ::matx::detail::log_at(5, __VA_ARGS__)

namespace matx {

// ---------------------------------------------------------------- executor
class HIPExecutor {
 public:
  using hip_executor = bool;
  using matx_executor = bool;
  explicit HIPExecutor(hipStream_t s = nullptr) : stream_(s) {}
  hipStream_t getStream() const { return stream_; }
  void sync() const { detail::hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }
  // executor timers (include/matx/executors/cuda_executor_common.h:142-169)
  void start_timer() {
    ensure_events();
    detail::hip_check(hipEventRecord(ev_.get()[0], stream_), "hipEventRecord");
  }
  void stop_timer() {
    ensure_events();
    detail::hip_check(hipEventRecord(ev_.get()[1], stream_), "hipEventRecord");
  }
  float get_time_ms() {
    float ms = 0.f;
    detail::hip_check(hipEventSynchronize(ev_.get()[1]), "hipEventSynchronize");
    detail::hip_check(hipEventElapsedTime(&ms, ev_.get()[0], ev_.get()[1]), "hipEventElapsedTime");
    return ms;
  }

 private:
  void ensure_events() {
    if (!ev_) {
      ev_ = std::shared_ptr<hipEvent_t>(new hipEvent_t[2], [](hipEvent_t *e) {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
        delete[] e;
      });
      detail::hip_check(hipEventCreate(&ev_.get()[0]), "hipEventCreate");
      detail::hip_check(hipEventCreate(&ev_.get()[1]), "hipEventCreate");
    }
  }
  hipStream_t stream_;
  std::shared_ptr<hipEvent_t> ev_;
};

// ----------------------------------------------------------------- tensors
template <typename T, int RANK>
class tensor_t {
  static_assert(RANK >= 0 && RANK <= MXR_MAX_RANK, "rank");

 public:
  using value_type = T;
  using type = T;
  using scalar_type = T;
  static constexpr int Rank() { return RANK; }

  tensor_t() = default;
  // Allocates a contiguous tensor, in managed memory by default as MatX's
  // make_tensor does (include/matx/core/make_tensor.h:57-58).  Where the
  // device cannot migrate managed pages (no concurrent managed access, as
  // on an MI355X without XNACK), hipMallocManaged returns host memory that
  // kernels reach over the host link at under 100 GB/s; there the managed
  // default allocates device memory instead, and host code reads elements
  // through copies (host element access throws).
  explicit tensor_t(const std::array<index_t, RANK> &shape, matxMemorySpace_t space = MATX_MANAGED_MEMORY) {
    shape_ = shape;
    index_t s = 1;
    for (int d = RANK - 1; d >= 0; --d) {
      stride_[d] = s;
      s *= shape_[d];
    }
    const size_t bytes = sizeof(T) * (size_t)(s > 0 ? s : 1);
    void *p = nullptr;
    if (space == MATX_MANAGED_MEMORY && !detail::managed_memory_migrates()) space = MATX_DEVICE_MEMORY;
    if (space == MATX_MANAGED_MEMORY) detail::hip_check(hipMallocManaged(&p, bytes), "hipMallocManaged");
    else if (space == MATX_DEVICE_MEMORY) detail::hip_check(hipMalloc(&p, bytes), "hipMalloc");
    else detail::hip_check(hipHostMalloc(&p, bytes), "hipHostMalloc");
    const bool host = space == MATX_HOST_MALLOC_MEMORY;
    storage_ = std::shared_ptr<void>(p, [host](void *q) {
      if (host) (void)hipHostFree(q);
      else (void)hipFree(q);
    });
    data_ = static_cast<T *>(p);
    space_ = space;
  }
  tensor_t(std::shared_ptr<void> st, T *data, const std::array<index_t, RANK> &shape,
           const std::array<index_t, RANK> &stride, matxMemorySpace_t space)
      : storage_(std::move(st)), data_(data), shape_(shape), stride_(stride), space_(space) {}

  MATX_AMD_HD index_t Size(int d) const { return shape_[d]; }
  MATX_AMD_HD index_t Stride(int d) const { return stride_[d]; }
  MATX_AMD_HD T *Data() const { return data_; }
  const std::array<index_t, RANK> &Shape() const { return shape_; }
  const std::array<index_t, RANK> &Strides() const { return stride_; }
  const std::shared_ptr<void> &Storage() const { return storage_; }
  matxMemorySpace_t Space() const { return space_; }
  index_t TotalSize() const {
    index_t n = 1;
    for (int d = 0; d < RANK; ++d) n *= shape_[d];
    return n;
  }
  bool IsContiguous() const {
    index_t s = 1;
    for (int d = RANK - 1; d >= 0; --d) {
      if (shape_[d] != 1 && stride_[d] != s) return false;
      s *= shape_[d];
    }
    return true;
  }
  // element access: on the host for managed or host memory (synchronise
  // first); in device code of a user-defined operator (BaseOp) for any
  // device-visible memory
  template <typename... I>
  MATX_AMD_HD T &operator()(I... idx) const {
    static_assert(sizeof...(I) == RANK, "index count");
#if !defined(__HIP_DEVICE_COMPILE__)
    if (space_ == MATX_DEVICE_MEMORY)
      throw matxException(MXR_INVALID_PARAMETER, "host access to an element of a tensor in device memory");
#endif
    const index_t ix[] = {(index_t)idx..., 0};
    index_t off = 0;
    for (int d = 0; d < RANK; ++d) off += ix[d] * stride_[d];
    return data_[off];
  }
  // view with permuted dimensions (include/matx/core/tensor.h:894-921)
  tensor_t Permute(const std::array<int, RANK> &dims) const {
    std::array<index_t, RANK> sh, st;
    for (int d = 0; d < RANK; ++d) {
      sh[d] = shape_[dims[d]];
      st[d] = stride_[dims[d]];
    }
    return tensor_t(storage_, data_, sh, st, space_);
  }
  // broadcast view of higher rank: matxKeepDim keeps this tensor's dims in
  // order, other entries are new zero-stride dims (tensor.h:1101-1114)
  template <int N>
  tensor_t<T, N> Clone(const std::array<index_t, N> &shape) const {
    std::array<index_t, N> sh, st;
    int k = 0;
    for (int d = 0; d < N; ++d) {
      if (shape[d] == matxKeepDim) {
        sh[d] = shape_[k];
        st[d] = stride_[k];
        ++k;
      } else {
        sh[d] = shape[d];
        st[d] = 0;
      }
    }
    if (k != RANK) throw matxException(MXR_INVALID_DIM, "Clone: matxKeepDim count != rank");
    return tensor_t<T, N>(storage_, data_, sh, st, space_);
  }
  void SetVals(std::initializer_list<T> v) {
    static_assert(RANK == 1, "SetVals(list) on rank 1");
    std::vector<T> h(v);
    put_host(h);
  }
  void SetVals(std::initializer_list<std::initializer_list<T>> v) {
    static_assert(RANK == 2, "SetVals(list of lists) on rank 2");
    std::vector<T> h;
    for (auto &r : v) h.insert(h.end(), r.begin(), r.end());
    put_host(h);
  }
  void PrefetchDevice(hipStream_t s) const {
    if (space_ != MATX_MANAGED_MEMORY) return;
    int dev = 0;
    detail::hip_check(hipGetDevice(&dev), "hipGetDevice");
    (void)hipMemPrefetchAsync(data_, sizeof(T) * (size_t)TotalSize(), dev, s);
  }
  mxr_tensor desc() const {
    mxr_tensor t{};
    t.data = (void *)data_;
    t.dtype = detail::dtype_of<T>::value;
    if (RANK == 0) {  // a scalar travels as a one-element vector
      t.rank = 1;
      t.shape[0] = 1;
      t.stride[0] = 1;
      return t;
    }
    t.rank = RANK;
    for (int d = 0; d < RANK; ++d) {
      t.shape[d] = shape_[d];
      t.stride[d] = stride_[d];
    }
    return t;
  }

  template <typename Op>
  auto operator=(const Op &op) const;

  // DLPack export (include/matx/core/tensor.h:1471-1555): a DLManagedTensor
  // of this view, no copy, that holds a reference to the storage until the
  // consumer calls its deleter.  Device memory and managed memory export as
  // ROCm device memory, pinned host memory as ROCm host memory.
  mxr_dl_managed *ToDlPack() const {
    int dev = 0;
    detail::hip_check(hipGetDevice(&dev), "hipGetDevice");
    auto *hold = new std::shared_ptr<void>(storage_);
    mxr_tensor d = desc();
    if (RANK == 0) d.rank = 0;
    mxr_dl_managed *m = nullptr;
    const mxr_status st = mxr_tensor_to_dlpack(
        &d, space_ == MATX_HOST_MALLOC_MEMORY ? 11 : 10, dev,
        [](void *p) { delete static_cast<std::shared_ptr<void> *>(p); }, hold, &m);
    if (st != MXR_SUCCESS) {
      delete hold;
      detail::check(st);
    }
    return m;
  }

 private:
  void put_host(const std::vector<T> &h) {
    if ((index_t)h.size() != TotalSize() || !IsContiguous())
      throw matxException(MXR_INVALID_SIZE, "SetVals: value count does not match the tensor");
    detail::hip_check(hipMemcpy(data_, h.data(), sizeof(T) * h.size(), hipMemcpyDefault), "hipMemcpy");
  }
  std::shared_ptr<void> storage_;
  T *data_ = nullptr;
  std::array<index_t, RANK> shape_{}, stride_{};
  matxMemorySpace_t space_ = MATX_MANAGED_MEMORY;
};

template <typename T, int RANK>
void make_tensor(tensor_t<T, RANK> &t, const std::array<index_t, RANK> &shape,
                 matxMemorySpace_t space = MATX_MANAGED_MEMORY) {
  t = tensor_t<T, RANK>(shape, space);
}
template <typename T>
void make_tensor(tensor_t<T, 0> &t, matxMemorySpace_t space = MATX_MANAGED_MEMORY) {
  t = tensor_t<T, 0>(std::array<index_t, 0>{}, space);
}
template <typename T, int RANK>
tensor_t<T, RANK> make_tensor(const std::array<index_t, RANK> &shape,
                              matxMemorySpace_t space = MATX_MANAGED_MEMORY) {
  return tensor_t<T, RANK>(shape, space);
}

// slice(t, starts, ends): matxEnd means "to the end", matxDropDim in ends
// removes the dimension (OutRank < RANK) (include/matx/operators/slice.h).
template <int OUT = -1, typename T, int RANK>
auto slice(const tensor_t<T, RANK> &t, const std::array<index_t, RANK> &starts,
           const std::array<index_t, RANK> &ends) {
  constexpr int R = OUT < 0 ? RANK : OUT;
  std::array<index_t, R> sh{}, st{};
  index_t off = 0;
  int k = 0;
  for (int d = 0; d < RANK; ++d) {
    const index_t b = starts[d];
    off += b * t.Stride(d);
    if (ends[d] == matxDropDim) continue;
    const index_t e = ends[d] == matxEnd ? t.Size(d) : ends[d];
    if (b < 0 || e > t.Size(d) || e < b) throw matxException(MXR_INVALID_SIZE, "slice out of range");
    if (k >= R) throw matxException(MXR_INVALID_DIM, "slice: output rank mismatch");
    sh[k] = e - b;
    st[k] = t.Stride(d);
    ++k;
  }
  if (k != R) throw matxException(MXR_INVALID_DIM, "slice: output rank mismatch");
  return tensor_t<T, R>(t.Storage(), t.Data() + off, sh, st, t.Space());
}

// --------------------------------------------------- elementwise expressions
namespace detail {
struct Node;
using NodeP = std::shared_ptr<const Node>;
struct Node {
  enum Kind { TENSOR, CONST, GEN, OP } kind;
  mxr_tensor t{};                 // TENSOR
  std::shared_ptr<void> keep;     // TENSOR storage lifetime
  cfloat c{};                     // CONST
  int op = 0, a = 0, b = 0;       // GEN (op, dim a, length b) or OP
  std::vector<NodeP> args;        // OP
  std::vector<index_t> shape;     // GEN / CONST from a shaped generator (ones, hamming)
};
}  // namespace detail

// An elementwise expression (rank given by the output it is assigned to).
struct Expr {
  detail::NodeP n;
};

namespace detail {
template <typename X>
Expr to_expr(const X &x) {
  if constexpr (std::is_same_v<X, Expr>) {
    return x;
  } else if constexpr (std::is_arithmetic_v<X>) {
    auto n = std::make_shared<Node>();
    n->kind = Node::CONST;
    n->c = cfloat((float)x, 0.f);
    return Expr{n};
  } else if constexpr (std::is_same_v<X, cfloat>) {
    auto n = std::make_shared<Node>();
    n->kind = Node::CONST;
    n->c = x;
    return Expr{n};
  } else {
    auto n = std::make_shared<Node>();
    n->kind = Node::TENSOR;
    n->t = x.desc();
    n->keep = x.Storage();
    return Expr{n};
  }
}
template <typename X>
struct is_exprish : std::false_type {};
template <>
struct is_exprish<Expr> : std::true_type {};
template <typename T, int R>
struct is_exprish<tensor_t<T, R>> : std::true_type {};
template <typename A, typename B>
constexpr bool binop_ok = (is_exprish<A>::value || is_exprish<B>::value) &&
                          (is_exprish<A>::value || std::is_arithmetic_v<A> || std::is_same_v<A, cfloat>) &&
                          (is_exprish<B>::value || std::is_arithmetic_v<B> || std::is_same_v<B, cfloat>);
inline Expr mk(int op, std::vector<Expr> a) {
  auto n = std::make_shared<Node>();
  n->kind = Node::OP;
  n->op = op;
  for (auto &e : a) n->args.push_back(e.n);
  return Expr{n};
}
// Shape of an expression: its tensor leaves' and shaped generators' extents
// broadcast against each other (aligned at the last dimension, extent 1
// stretches), as MatX derives an operator's Size() from its operands.
// Empty when nothing in the expression has a shape.
inline void merge_shape(std::vector<index_t> &acc, const std::vector<index_t> &sh) {
  if (acc.size() < sh.size()) acc.insert(acc.begin(), sh.size() - acc.size(), 1);
  const size_t pad = acc.size() - sh.size();
  for (size_t d = 0; d < sh.size(); ++d) {
    index_t &a = acc[pad + d];
    if (a == sh[d] || sh[d] == 1) continue;
    if (a != 1) throw matxException(MXR_INVALID_SIZE, "expression operands have incompatible shapes");
    a = sh[d];
  }
}
inline void expr_shape_into(const NodeP &n, std::vector<index_t> &acc, bool &any) {
  if (n->kind == Node::TENSOR) {
    merge_shape(acc, std::vector<index_t>(n->t.shape, n->t.shape + n->t.rank));
    any = true;
  } else if (!n->shape.empty()) {
    merge_shape(acc, n->shape);
    any = true;
  }
  for (auto &a : n->args) expr_shape_into(a, acc, any);
}
inline std::vector<index_t> expr_shape(const Expr &e) {
  std::vector<index_t> acc;
  bool any = false;
  expr_shape_into(e.n, acc, any);
  return any ? acc : std::vector<index_t>{};
}
}  // namespace detail

#define MATX_AMD_BINOP(OPSYM, CODE)                                                      \
  template <typename A, typename B, std::enable_if_t<detail::binop_ok<A, B>, int> = 0> \
  Expr operator OPSYM(const A &a, const B &b) {                                        \
    return detail::mk(CODE, {detail::to_expr(a), detail::to_expr(b)});                 \
  }
MATX_AMD_BINOP(+, MXR_EW_ADD)
MATX_AMD_BINOP(-, MXR_EW_SUB)
MATX_AMD_BINOP(*, MXR_EW_MUL)
MATX_AMD_BINOP(/, MXR_EW_DIV)
MATX_AMD_BINOP(>, MXR_EW_GT)
MATX_AMD_BINOP(<, MXR_EW_LT)
MATX_AMD_BINOP(>=, MXR_EW_GE)
MATX_AMD_BINOP(<=, MXR_EW_LE)
#undef MATX_AMD_BINOP

template <typename A>
Expr conj(const A &a) { return detail::mk(MXR_EW_CONJ, {detail::to_expr(a)}); }
template <typename A>
Expr abs2(const A &a) { return detail::mk(MXR_EW_ABS2, {detail::to_expr(a)}); }
template <typename A>
Expr abs(const A &a) { return detail::mk(MXR_EW_ABS, {detail::to_expr(a)}); }
template <typename A>
Expr sqrt(const A &a) { return detail::mk(MXR_EW_SQRT, {detail::to_expr(a)}); }
template <typename A>
Expr exp(const A &a) { return detail::mk(MXR_EW_EXP, {detail::to_expr(a)}); }
template <typename A, typename B>
Expr pow(const A &a, const B &b) { return detail::mk(MXR_EW_POW, {detail::to_expr(a), detail::to_expr(b)}); }
// IF(cond) a else b
template <typename C, typename A, typename B>
Expr select(const C &c, const A &a, const B &b) {
  return detail::mk(MXR_EW_SELECT, {detail::to_expr(a), detail::to_expr(b), detail::to_expr(c)});
}
// hamming<Dim>(shape): 0.54 - 0.46 cos(2 pi i / (n-1)) along Dim, n = shape[Dim]
// (include/matx/generators/hamming.h:146-178)
template <int Dim, int Rank = 1, typename T = float>
Expr hamming(const std::array<index_t, Rank> &shape) {
  auto n = std::make_shared<detail::Node>();
  n->kind = detail::Node::GEN;
  n->op = MXR_EW_HAMMING;
  n->a = Dim;
  n->b = (int)shape[Dim];
  n->shape.assign(shape.begin(), shape.end());
  return Expr{n};
}
// ones(shape) / zeros(shape) (include/matx/generators/ones.h, zeros.h): a
// constant that carries its shape, so a transform can take it as an operand
// (conv2d(ones({C, P, R}), mask, FULL)); the rank comes from the braced list.
template <typename T = float, int Rank>
Expr ones(const index_t (&shape)[Rank]) {
  Expr e = detail::to_expr(1.0f);
  auto n = std::make_shared<detail::Node>(*e.n);
  n->shape.assign(shape, shape + Rank);
  return Expr{n};
}
template <typename T = float, int Rank>
Expr zeros(const index_t (&shape)[Rank]) {
  Expr e = detail::to_expr(0.0f);
  auto n = std::make_shared<detail::Node>(*e.n);
  n->shape.assign(shape, shape + Rank);
  return Expr{n};
}

// ------------------------------------------------------------ interop
// DLPack import (make_tensor(DLManagedTensor*), include/matx/core/
// make_tensor.h:1136-1190): the tensor takes ownership and runs the
// producer's deleter when its last view goes.  mxr_dl_managed (mxr.h) has
// the layout of DLPack's DLManagedTensor.
template <typename T, int RANK>
tensor_t<T, RANK> make_tensor_from_dlpack(mxr_dl_managed *m) {
  mxr_tensor d{};
  detail::check(mxr_tensor_from_dlpack(&d, &m->dl_tensor));
  if (d.rank != RANK || d.dtype != detail::dtype_of<T>::value)
    throw matxException(MXR_INVALID_TYPE, "make_tensor_from_dlpack: rank or element type differs");
  std::array<index_t, RANK> sh{}, st{};
  for (int i = 0; i < RANK; ++i) {
    sh[i] = d.shape[i];
    st[i] = d.stride[i];
  }
  std::shared_ptr<void> own(m, [](void *p) {
    auto *mm = static_cast<mxr_dl_managed *>(p);
    if (mm->deleter) mm->deleter(mm);
  });
  return tensor_t<T, RANK>(own, static_cast<T *>(d.data), sh, st, MATX_DEVICE_MEMORY);
}

namespace io {
namespace npy_detail {
template <typename T>
inline const char *descr() {
  if constexpr (std::is_same_v<T, cfloat>) return "<c8";
  else if constexpr (std::is_same_v<T, float>) return "<f4";
  else if constexpr (std::is_same_v<T, int>) return "<i4";
  else static_assert(sizeof(T) == 0, "npy: complex<float>, float or int elements");
}
template <typename T, int RANK>
void require_contiguous(const tensor_t<T, RANK> &t, const char *fn) {
  if (!t.IsContiguous()) throw matxException(MXR_INVALID_PARAMETER, std::string(fn) + ": view is not contiguous");
}
}  // namespace npy_detail

// write_npy / read_npy (include/matx/file_io/file_io.h:321-362), with the
// .npy format written and parsed here (version 1.0 written; 1.0, 2.0 and 3.0
// read; little-endian complex<float>, float or int; C order).  read_npy fills
// an existing contiguous view whose shape must match the file's.
template <typename T, int RANK>
void write_npy(const tensor_t<T, RANK> &t, const std::string &fname) {
  npy_detail::require_contiguous(t, "write_npy");
  std::string shape = "(";
  for (int d = 0; d < RANK; ++d) shape += std::to_string(t.Size(d)) + (RANK == 1 ? "," : (d + 1 < RANK ? ", " : ""));
  shape += ")";
  std::string hdr = std::string("{'descr': '") + npy_detail::descr<T>() + "', 'fortran_order': False, 'shape': " +
                    shape + ", }";
  const size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr.push_back('\n');
  std::vector<T> h((size_t)t.TotalSize());
  detail::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (!h.empty())
    detail::hip_check(hipMemcpy(h.data(), t.Data(), sizeof(T) * h.size(), hipMemcpyDefault), "copy");
  FILE *f = fopen(fname.c_str(), "wb");
  if (!f) throw matxException(MXR_INVALID_PARAMETER, "write_npy: cannot open " + fname);
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t hl = (uint16_t)hdr.size();
  const bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&hl, 2, 1, f) == 1 &&
                  fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
                  (h.empty() || fwrite(h.data(), sizeof(T), h.size(), f) == h.size());
  fclose(f);
  if (!ok) throw matxException(MXR_INVALID_PARAMETER, "write_npy: short write to " + fname);
}

template <typename T, int RANK>
void read_npy(tensor_t<T, RANK> &t, const std::string &fname) {
  npy_detail::require_contiguous(t, "read_npy");
  FILE *f = fopen(fname.c_str(), "rb");
  if (!f) throw matxException(MXR_INVALID_PARAMETER, "read_npy: failed to read [" + fname + "], does not exist");
  auto fail = [&](const std::string &why) {
    fclose(f);
    throw matxException(MXR_INVALID_PARAMETER, "read_npy: " + fname + ": " + why);
  };
  unsigned char magic[8];
  if (fread(magic, 1, 8, f) != 8 || magic[0] != 0x93 || std::memcmp(magic + 1, "NUMPY", 5) != 0)
    fail("not an .npy file");
  uint32_t hl = 0;
  if (magic[6] == 1) {
    uint16_t h16 = 0;
    if (fread(&h16, 2, 1, f) != 1) fail("truncated header");
    hl = h16;
  } else if (magic[6] == 2 || magic[6] == 3) {
    if (fread(&hl, 4, 1, f) != 1) fail("truncated header");
  } else {
    fail("unknown format version");
  }
  std::string hdr(hl, '\0');
  if (fread(&hdr[0], 1, hl, f) != hl) fail("truncated header");
  auto field = [&](const char *key) -> std::string {
    const size_t k = hdr.find(key);
    if (k == std::string::npos) fail(std::string("no ") + key);
    size_t v = hdr.find(':', k) + 1;
    while (v < hdr.size() && hdr[v] == ' ') ++v;
    return hdr.substr(v);
  };
  const std::string ds = field("'descr'");
  const std::string want = std::string("'") + npy_detail::descr<T>() + "'";
  if (ds.compare(0, want.size(), want) != 0) fail("element type is not " + want);
  if (field("'fortran_order'").compare(0, 5, "False") != 0) fail("Fortran order is not supported");
  const std::string sh = field("'shape'");
  std::vector<index_t> dims;
  for (size_t i = 1; i < sh.size() && sh[i] != ')';) {
    if (std::isdigit((unsigned char)sh[i])) {
      size_t e = i;
      while (e < sh.size() && std::isdigit((unsigned char)sh[e])) ++e;
      dims.push_back((index_t)std::stoll(sh.substr(i, e - i)));
      i = e;
    } else {
      ++i;
    }
  }
  if ((int)dims.size() != RANK) fail("rank differs from the tensor's");
  for (int d = 0; d < RANK; ++d)
    if (dims[d] != t.Size(d)) fail("shape differs from the tensor's");
  std::vector<T> h((size_t)t.TotalSize());
  if (!h.empty() && fread(h.data(), sizeof(T), h.size(), f) != h.size()) fail("truncated data");
  fclose(f);
  if (!h.empty())
    detail::hip_check(hipMemcpy(t.Data(), h.data(), sizeof(T) * h.size(), hipMemcpyDefault), "copy");
}
}  // namespace io

// --------------------------------------------------------- transform ops
namespace detail {
struct FFTSpec {
  Expr in;
  std::vector<index_t> in_shape;  // shape of `in` when it must be materialised
  index_t n = 0;
  int dir = MXR_FFT_FORWARD;
  int norm = (int)FFTNorm::BACKWARD;
  int axis = -1;
  int kind = 0;  // 0 complex-to-complex, 1 real-to-complex, 2 complex-to-real
};
}  // namespace detail

struct FFTOp {
  detail::FFTSpec s;
};
// 2-D transform over the last two axes (fft2 / ifft2 / rfft2 / irfft2)
struct FFT2Op {
  detail::FFTSpec s;
};
struct FFTConvOp {
  Expr x, w;
  index_t n;
  int conj_w;
};
struct Conv1DOp {
  Expr in, f;
  int mode;
  int method = MATX_C_METHOD_DIRECT;
  int corr = 0;  // correlate: conv1d with reverse(conj(f)) (FFT kernel)
};
struct Conv2DOp {
  Expr in, f;
  int mode;
};
struct SumOp {
  Expr in;
  std::vector<index_t> in_shape;
};

template <typename T, int R>
FFTOp fft(const tensor_t<T, R> &a, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_FORWARD, (int)norm, R - 1}};
}
template <typename T, int R>
FFTOp ifft(const tensor_t<T, R> &a, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_INVERSE, (int)norm, R - 1}};
}
// fft along an explicit axis: fft(a, {axis}) (include/matx/operators/fft.h:565-577)
template <typename T, int R>
FFTOp fft(const tensor_t<T, R> &a, std::array<int, 1> axis, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_FORWARD, (int)norm, axis[0]}};
}
template <typename T, int R>
FFTOp ifft(const tensor_t<T, R> &a, std::array<int, 1> axis, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_INVERSE, (int)norm, axis[0]}};
}
// rfft / irfft (include/matx/operators/fft.h:580-740): real input, n/2 + 1
// bins out; bins in, n real samples out (n = 0: 2 (N - 1))
template <typename T, int R>
FFTOp rfft(const tensor_t<T, R> &a, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_FORWARD, (int)norm, R - 1, 1}};
}
template <typename T, int R>
FFTOp irfft(const tensor_t<T, R> &a, index_t n = 0, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFTOp{{detail::to_expr(a), {}, n, MXR_FFT_INVERSE, (int)norm, R - 1, 2}};
}
// fft2 / ifft2 / rfft2 / irfft2 over the last two axes (include/matx/operators/fft.h:743-1310)
template <typename T, int R>
FFT2Op fft2(const tensor_t<T, R> &a, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFT2Op{{detail::to_expr(a), {}, 0, MXR_FFT_FORWARD, (int)norm, R - 1, 0}};
}
template <typename T, int R>
FFT2Op ifft2(const tensor_t<T, R> &a, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFT2Op{{detail::to_expr(a), {}, 0, MXR_FFT_INVERSE, (int)norm, R - 1, 0}};
}
template <typename T, int R>
FFT2Op rfft2(const tensor_t<T, R> &a, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFT2Op{{detail::to_expr(a), {}, 0, MXR_FFT_FORWARD, (int)norm, R - 1, 1}};
}
// out's last extent gives the real length n
template <typename T, int R>
FFT2Op irfft2(const tensor_t<T, R> &a, FFTNorm norm = FFTNorm::BACKWARD) {
  return FFT2Op{{detail::to_expr(a), {}, 0, MXR_FFT_INVERSE, (int)norm, R - 1, 2}};
}
// out = ifft(fft(x, n) * conj(fft(w, n)))  (circular cross-correlation)
template <typename X, typename W>
FFTConvOp fft_conv(const X &x, const W &w, index_t n, bool conj_w = true) {
  return FFTConvOp{detail::to_expr(x), detail::to_expr(w), n, conj_w ? 1 : 0};
}
template <typename X, typename F>
Conv1DOp conv1d(const X &x, const F &f, matxConvCorrMode_t mode,
                matxConvCorrMethod_t method = MATX_C_METHOD_DIRECT) {
  return Conv1DOp{detail::to_expr(x), detail::to_expr(f), (int)mode, (int)method, 0};
}
// corr(x, f) = conv1d(x, reverse(conj(f))) (include/matx/transforms/corr.h:60-67);
// both methods run mxr_conv1d_fft, whose spectrum kernel reverses the filter.
template <typename X, typename F>
Conv1DOp corr(const X &x, const F &f, matxConvCorrMode_t mode,
              matxConvCorrMethod_t method = MATX_C_METHOD_DIRECT) {
  (void)method;
  return Conv1DOp{detail::to_expr(x), detail::to_expr(f), (int)mode, MATX_C_METHOD_FFT, 1};
}
template <typename X, typename F>
Conv2DOp conv2d(const X &x, const F &f, matxConvCorrMode_t mode) {
  return Conv2DOp{detail::to_expr(x), detail::to_expr(f), (int)mode};
}
template <typename T, int R>
SumOp sum(const tensor_t<T, R> &a) {
  return SumOp{detail::to_expr(a), std::vector<index_t>(a.Shape().begin(), a.Shape().end())};
}
// sum(expr): the reduction runs over the expression's shape, taken from its
// operands (sum(abs2(waveformPart)), examples/simple_radar_pipeline.h:291)
inline SumOp sum(const Expr &e) {
  auto sh = detail::expr_shape(e);
  if (sh.empty()) throw matxException(MXR_INVALID_SIZE, "sum: the expression has no shaped operand");
  return SumOp{e, std::move(sh)};
}

// ---------------------------------------------------------------- run()
namespace detail {
struct Lowered {
  std::vector<mxr_ew_instr> prog;
  std::vector<mxr_tensor> ins;
  std::vector<float> consts;
};
// Aligns an input view to the output's trailing dims (zero strides elsewhere).
inline mxr_tensor broadcast_to(const mxr_tensor &in, const mxr_tensor &out) {
  mxr_tensor t = in;
  t.rank = out.rank;
  const int pad = out.rank - in.rank;
  if (pad < 0) throw matxException(MXR_INVALID_DIM, "expression input has a higher rank than its output");
  for (int d = out.rank - 1; d >= 0; --d) {
    const int s = d - pad;
    if (s >= 0 && in.shape[s] == out.shape[d]) {
      t.shape[d] = in.shape[s];
      t.stride[d] = in.stride[s];
    } else if (s < 0 || in.shape[s] == 1) {
      t.shape[d] = out.shape[d];
      t.stride[d] = 0;
    } else {
      throw matxException(MXR_INVALID_SIZE, "expression input extent does not match the output");
    }
  }
  return t;
}
inline void lower(const NodeP &n, const mxr_tensor &out, Lowered &L) {
  switch (n->kind) {
    case Node::TENSOR: {
      L.ins.push_back(broadcast_to(n->t, out));
      L.prog.push_back({MXR_EW_LOAD, (int)L.ins.size() - 1, 0});
      break;
    }
    case Node::CONST:
      L.consts.push_back(n->c.real());
      L.consts.push_back(n->c.imag());
      L.prog.push_back({MXR_EW_CONST, (int)L.consts.size() / 2 - 1, 0});
      break;
    case Node::GEN: {
      // generator dims count from the output's leading dimension
      L.prog.push_back({n->op, n->a, n->b});
      break;
    }
    case Node::OP:
      for (auto &a : n->args) lower(a, out, L);
      L.prog.push_back({n->op, 0, 0});
      break;
  }
}
inline void eval_into(const mxr_tensor &out, const Expr &e, hipStream_t s) {
  Lowered L;
  lower(e.n, out, L);
  std::vector<const mxr_tensor *> ptrs;
  for (auto &t : L.ins) ptrs.push_back(&t);
  check(mxr_ew_eval(&out, ptrs.data(), (int)ptrs.size(), L.prog.data(), (int)L.prog.size(),
                    L.consts.empty() ? nullptr : L.consts.data(), (int)L.consts.size() / 2, s));
}
// A transform operand: a tensor view as is, or an expression materialised
// into a temporary of the given shape and dtype.
struct Operand {
  mxr_tensor t{};
  std::shared_ptr<void> tmp;
};
inline Operand operand(const Expr &e, const std::vector<index_t> &shape_in, int dtype, hipStream_t s) {
  Operand o;
  if (e.n->kind == Node::TENSOR) {
    o.t = e.n->t;
    return o;
  }
  const std::vector<index_t> shape = shape_in.empty() ? expr_shape(e) : shape_in;
  if (shape.empty()) throw matxException(MXR_INVALID_PARAMETER, "expression operand has no shape");
  index_t total = 1;
  mxr_tensor t{};
  t.rank = (int)shape.size();
  t.dtype = dtype;
  for (int d = t.rank - 1; d >= 0; --d) {
    t.shape[d] = shape[d];
    t.stride[d] = total;
    total *= shape[d];
  }
  void *p = nullptr;
  hip_check(hipMallocAsync(&p, (dtype == MXR_C64 ? 8 : 4) * (size_t)total, s), "hipMallocAsync");
  o.tmp = std::shared_ptr<void>(p, [s](void *q) { (void)hipFreeAsync(q, s); });
  t.data = p;
  eval_into(t, e, s);
  o.t = t;
  return o;
}
}  // namespace detail

template <typename L, typename Op>
struct SetOp {
  L lhs;
  Op op;
  void run(const HIPExecutor &ex) const { run(ex.getStream()); }
  void run(hipStream_t s = nullptr) const {
    const mxr_tensor out = lhs.desc();
    using namespace detail;
    if constexpr (std::is_same_v<Op, Expr>) {
      eval_into(out, op, s);
    } else if constexpr (std::is_same_v<Op, FFTOp>) {
      auto in = operand(op.s.in, op.s.in_shape, op.s.kind == 1 ? MXR_F32 : MXR_C64, s);
      const int axis = op.s.axis < 0 ? out.rank - 1 : op.s.axis;
      if (op.s.kind == 1) check(mxr_fft_r2c(&out, &in.t, op.s.n, op.s.norm, axis, s));
      else if (op.s.kind == 2) check(mxr_fft_c2r(&out, &in.t, op.s.n, op.s.norm, axis, s));
      else check(mxr_fft_c2c(&out, &in.t, op.s.n, op.s.dir, op.s.norm, axis, s));
    } else if constexpr (std::is_same_v<Op, FFT2Op>) {
      auto in = operand(op.s.in, op.s.in_shape, op.s.kind == 1 ? MXR_F32 : MXR_C64, s);
      const int r = out.rank;
      if (op.s.kind == 0) {  // rows, then columns in place
        check(mxr_fft_c2c(&out, &in.t, 0, op.s.dir, op.s.norm, r - 1, s));
        check(mxr_fft_c2c(&out, &out, 0, op.s.dir, op.s.norm, r - 2, s));
      } else if (op.s.kind == 1) {
        check(mxr_fft_r2c(&out, &in.t, 0, op.s.norm, r - 1, s));
        check(mxr_fft_c2c(&out, &out, 0, MXR_FFT_FORWARD, op.s.norm, r - 2, s));
      } else {  // columns into a stream-ordered temporary, then the real rows
        mxr_tensor t{};
        t.rank = in.t.rank;
        t.dtype = MXR_C64;
        index_t total = 1;
        for (int d = t.rank - 1; d >= 0; --d) {
          t.shape[d] = in.t.shape[d];
          t.stride[d] = total;
          total *= t.shape[d];
        }
        void *p = nullptr;
        hip_check(hipMallocAsync(&p, 8 * (size_t)total, s), "hipMallocAsync");
        std::shared_ptr<void> hold(p, [s](void *q) { (void)hipFreeAsync(q, s); });
        t.data = p;
        check(mxr_fft_c2c(&t, &in.t, 0, MXR_FFT_INVERSE, op.s.norm, r - 2, s));
        check(mxr_fft_c2r(&out, &t, out.shape[r - 1], op.s.norm, r - 1, s));
      }
    } else if constexpr (std::is_same_v<Op, FFTConvOp>) {
      auto x = operand(op.x, {}, MXR_C64, s);
      auto w = operand(op.w, {}, MXR_C64, s);
      check(mxr_fft_conv(&out, &x.t, &w.t, op.n, op.conj_w, s));
    } else if constexpr (std::is_same_v<Op, Conv1DOp>) {
      auto x = operand(op.in, {}, MXR_C64, s);
      auto f = operand(op.f, {}, MXR_F32, s);
      if (op.method == MATX_C_METHOD_FFT || op.corr)
        check(mxr_conv1d_fft(&out, &x.t, &f.t, op.mode, op.corr, s));
      else
        check(mxr_conv1d(&out, &x.t, &f.t, op.mode, s));
    } else if constexpr (std::is_same_v<Op, Conv2DOp>) {
      auto x = operand(op.in, {}, MXR_F32, s);
      auto f = operand(op.f, {}, MXR_F32, s);
      check(mxr_conv2d(&out, &x.t, &f.t, op.mode, s));
    } else if constexpr (std::is_same_v<Op, SumOp>) {
      auto x = operand(op.in, op.in_shape, out.dtype, s);
      check(mxr_sum(&out, &x.t, s));
    }
  }
};

template <typename T, int RANK>
template <typename Op>
auto tensor_t<T, RANK>::operator=(const Op &op) const {
  if constexpr (std::is_same_v<Op, FFTOp> || std::is_same_v<Op, FFT2Op> || std::is_same_v<Op, FFTConvOp> ||
                std::is_same_v<Op, Conv1DOp> || std::is_same_v<Op, Conv2DOp> ||
                std::is_same_v<Op, SumOp>) {
    return SetOp<tensor_t, Op>{*this, op};
  } else {
    return SetOp<tensor_t, Expr>{*this, detail::to_expr(op)};
  }
}

// ------------------------------------------------ user-defined operators
// MatX's custom-operator form (examples/simple_radar_pipeline.h:50-150): a
// class derived from BaseOp<Op> with a templated __device__ operator() over
// the output's indices, Size(i) and Rank(); `Op(...).run(exec)` evaluates it
// once per output element.  The operator is copied byte for byte into the
// kernel's argument block (its tensor members are views: pointer, shape,
// strides), so the device code never constructs or destroys one.
namespace detail {
// Capability vocabulary user operators query in get_capability
// (include/matx/core/capabilities.h).  This launcher runs one element per
// thread, so CapType::ept is always ONE.
enum class ElementsPerThread : int {
  INVALID = 0, ONE = 1, TWO = 2, FOUR = 4, EIGHT = 8, SIXTEEN = 16, THIRTY_TWO = 32
};
enum class OperatorCapability : int { NONE = 0, ELEMENTS_PER_THREAD, ALIASED_MEMORY, SUPPORTS_JIT };
template <OperatorCapability Cap>
struct capability_attributes {
  using type = bool;
  static constexpr bool default_value = false;
};
template <>
struct capability_attributes<OperatorCapability::ELEMENTS_PER_THREAD> {
  using type = std::array<ElementsPerThread, 2>;  // {lowest, highest} supported
  static constexpr type default_value{ElementsPerThread::ONE, ElementsPerThread::THIRTY_TWO};
};
template <ElementsPerThread EPT, bool JIT = false>
struct CapabilityParams {
  static constexpr ElementsPerThread ept = EPT;
  static constexpr bool jit = JIT;
};
template <typename Op>
struct is_tensor_view : std::false_type {};
template <typename T, int R>
struct is_tensor_view<tensor_t<T, R>> : std::true_type {};
// Tensors and scalars have the defaults; an operator answers for itself.
template <OperatorCapability Cap, typename Op, typename In>
auto get_operator_capability(const Op &op, const In &in) {
  if constexpr (is_tensor_view<Op>::value || std::is_arithmetic_v<Op>) return capability_attributes<Cap>::default_value;
  else return op.template get_capability<Cap>(in);
}
// Element-per-thread ranges intersect; ALIASED_MEMORY is true if any operand
// aliases; every other flag holds only if it holds for all.
template <OperatorCapability Cap, typename A>
auto combine_capabilities(A a) { return a; }
template <OperatorCapability Cap, typename A, typename B, typename... Rest>
auto combine_capabilities(A a, B b, Rest... rest) {
  if constexpr (Cap == OperatorCapability::ELEMENTS_PER_THREAD) {
    A c{(int)a[0] > (int)b[0] ? a[0] : b[0], (int)a[1] < (int)b[1] ? a[1] : b[1]};
    return combine_capabilities<Cap>(c, rest...);
  } else if constexpr (Cap == OperatorCapability::ALIASED_MEMORY) {
    return combine_capabilities<Cap>(A(a || b), rest...);
  } else {
    return combine_capabilities<Cap>(A(a && b), rest...);
  }
}

#if defined(__HIPCC__)
template <typename Op>
struct alignas(alignof(Op)) OpImage {
  unsigned char b[sizeof(Op)];
};
template <int R>
struct OpExtents {
  index_t n[R > 0 ? R : 1];
};
template <typename Op, size_t... I>
__device__ __forceinline__ void call_user_op(Op &op, const index_t *ix, std::index_sequence<I...>) {
  op.template operator()<CapabilityParams<ElementsPerThread::ONE, false>>(ix[I]...);
}
template <typename Op, int R>
__global__ __launch_bounds__(256) void user_op_kernel(OpImage<Op> img, OpExtents<R> ext, index_t total) {
  index_t i = (index_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  index_t ix[R > 0 ? R : 1];
  for (int d = R - 1; d >= 0; --d) {
    ix[d] = i % ext.n[d];
    i /= ext.n[d];
  }
  Op &op = const_cast<Op &>(*reinterpret_cast<const Op *>(img.b));
  call_user_op(op, ix, std::make_index_sequence<(size_t)R>{});
}
template <typename Op>
void run_user_op(const Op &op, hipStream_t s) {
  constexpr int R = Op::Rank();
  OpExtents<R> ext{};
  index_t total = 1;
  for (int d = 0; d < R; ++d) {
    ext.n[d] = op.Size((uint32_t)d);
    total *= ext.n[d];
  }
  if (total <= 0) return;
  const index_t blocks = (total + 255) / 256;
  if (blocks > (index_t)std::numeric_limits<int32_t>::max())
    throw matxException(MXR_INVALID_SIZE, "user operator: output too large for one launch");
  OpImage<Op> img;
  std::memcpy(img.b, static_cast<const void *>(&op), sizeof(Op));
  user_op_kernel<Op, R><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(img, ext, total);
  hip_check(hipGetLastError(), "user operator launch");
}
#else
template <typename Op>
void run_user_op(const Op &, hipStream_t) {
  static_assert(sizeof(Op) == 0,
                "a user-defined operator (BaseOp) runs its own device code: compile this "
                "translation unit with hipcc");
}
#endif
}  // namespace detail

template <typename Derived>
class BaseOp {
 public:
  using matxop = bool;
  void run(const HIPExecutor &ex) const { run(ex.getStream()); }
  void run(hipStream_t s = nullptr) const { detail::run_user_op(static_cast<const Derived &>(*this), s); }
};

// --------------------------------------------------- fused radar stages
// One-call stage entries (pulse compression in place, canceller + window +
// Doppler FFT, CA-CFAR) for pipelines that do not need MatX's stage-by-stage
// intermediate buffers.
template <typename TY, typename TX, typename TW>
void pulse_compress(const TY &y, const TX &x, const TW &w, index_t n, int flags, const HIPExecutor &ex) {
  const mxr_tensor yd = y.desc(), xd = x.desc(), wd = w.desc();
  detail::check(mxr_pulse_compress(&yd, &xd, &wd, n, flags, ex.getStream()));
}

}  // namespace matx
