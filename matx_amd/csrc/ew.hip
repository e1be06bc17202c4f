// Fused elementwise expressions and the scalar sum.
//
// MatX fuses an elementwise expression tree into one templated kernel per
// (lhs = expr).run(exec) at compile time (include/matx/executors/kernel.h).
// Behind a C ABI the expression cannot be a C++ type, so the front end lowers
// it to a short postfix program.  By default the program is turned into HIP
// source and compiled once per expression shape (the expression compiler
// below); otherwise one precompiled kernel interprets it per element.  The program is the same for every lane, so its decode and
// branches run on the scalar unit; only the arithmetic is per lane.  Each
// lane handles EW_EPT consecutive elements of the innermost dimension, so the
// index decomposition is paid once per run of elements.
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdlib>
#include <chrono>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "mxr_cache.hpp"
#include "mxr_trace.hpp"
#include "mxr_internal.hpp"

namespace mxr {

constexpr int EW_MAXIN = 8, EW_MAXPROG = 48, EW_MAXCONST = 16, EW_STACK = 8;
constexpr int EW_EPT = 4;

struct EwIn {
  const void *p;
  int dtype;
  int64_t st[MXR_MAX_RANK];
};
struct EwArgs {
  void *out;
  int odtype;
  int rank;
  int64_t shape[MXR_MAX_RANK];
  int64_t ost[MXR_MAX_RANK];
  int64_t total_runs;  // runs of EW_EPT along the innermost dim
  int64_t runs_per_row;
  int nin, nprog;
  EwIn in[EW_MAXIN];
  mxr_ew_instr prog[EW_MAXPROG];
  float2 cst[EW_MAXCONST];
};
static_assert(sizeof(EwArgs) <= 4096, "the argument block must fit the kernel arguments");

__device__ __forceinline__ float2 ew_load(const EwIn &in, int64_t off) {
  switch (in.dtype) {
    case MXR_C64: return reinterpret_cast<const float2 *>(in.p)[off];
    case MXR_F32: return make_float2(reinterpret_cast<const float *>(in.p)[off], 0.f);
    default: return make_float2((float)reinterpret_cast<const int *>(in.p)[off], 0.f);
  }
}

__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {
  // MatX simplified complex division (core/defines.h:50-58 semantics)
  const float d = b.x * b.x + b.y * b.y;
  return make_float2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// The argument block (about 1.5 KB) travels by value in the kernel
// arguments, so a launch holds no pointer to host memory (and can be
// captured into a graph).
__global__ __launch_bounds__(256) void ew_kernel(const EwArgs a) {
  const int64_t run = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (run >= a.total_runs) return;
  const int r = a.rank;
  // coordinates of the first element of this run
  int64_t idx[MXR_MAX_RANK];
  int64_t rem = run / a.runs_per_row;
  idx[r - 1] = (run % a.runs_per_row) * EW_EPT;
  for (int d = r - 2; d >= 0; --d) {
    idx[d] = rem % a.shape[d];
    rem /= a.shape[d];
  }
  int64_t ooff = 0;
  int64_t ioff[EW_MAXIN];
  for (int k = 0; k < a.nin; ++k) ioff[k] = 0;
  for (int d = 0; d < r; ++d) {
    ooff += idx[d] * a.ost[d];
    for (int k = 0; k < a.nin; ++k) ioff[k] += idx[d] * a.in[k].st[d];
  }
  const int64_t last = idx[r - 1];
  const int cnt = (int)min((int64_t)EW_EPT, a.shape[r - 1] - last);
  for (int e = 0; e < cnt; ++e) {
    float2 stk[EW_STACK];
    int sp = 0;
    for (int pc = 0; pc < a.nprog; ++pc) {
      const mxr_ew_instr ins = a.prog[pc];
      switch (ins.op) {
        case MXR_EW_LOAD:
          stk[sp++] = ew_load(a.in[ins.a], ioff[ins.a] + e * a.in[ins.a].st[r - 1]);
          break;
        case MXR_EW_CONST: stk[sp++] = a.cst[ins.a]; break;
        case MXR_EW_HAMMING: {
          const int64_t i = ins.a == r - 1 ? last + e : idx[ins.a];
          stk[sp++] = make_float2(hamming_f32((int)i, ins.b), 0.f);
          break;
        }
        case MXR_EW_INDEX: {
          const int64_t i = ins.a == r - 1 ? last + e : idx[ins.a];
          stk[sp++] = make_float2((float)i, 0.f);
          break;
        }
        case MXR_EW_SELECT: {
          const float2 cnd = stk[--sp], y = stk[--sp], x = stk[--sp];
          stk[sp++] = cnd.x != 0.f ? x : y;
          break;
        }
        default: {
          if (ins.op >= MXR_EW_ADD && ins.op < MXR_EW_CONJ) {
            const float2 y = stk[--sp], x = stk[--sp];
            float2 z;
            switch (ins.op) {
              case MXR_EW_ADD: z = cadd(x, y); break;
              case MXR_EW_SUB: z = csub(x, y); break;
              case MXR_EW_MUL: z = cmul(x, y); break;
              case MXR_EW_DIV: z = (y.y == 0.f) ? make_float2(x.x / y.x, x.y / y.x) : cdiv(x, y); break;
              case MXR_EW_POW: z = make_float2(powf(x.x, y.x), 0.f); break;
              case MXR_EW_GT: z = make_float2(x.x > y.x ? 1.f : 0.f, 0.f); break;
              case MXR_EW_LT: z = make_float2(x.x < y.x ? 1.f : 0.f, 0.f); break;
              case MXR_EW_GE: z = make_float2(x.x >= y.x ? 1.f : 0.f, 0.f); break;
              case MXR_EW_LE: z = make_float2(x.x <= y.x ? 1.f : 0.f, 0.f); break;
              case MXR_EW_EQ: z = make_float2((x.x == y.x && x.y == y.y) ? 1.f : 0.f, 0.f); break;
              case MXR_EW_MAX: z = make_float2(fmaxf(x.x, y.x), 0.f); break;
              default: z = make_float2(fminf(x.x, y.x), 0.f); break;
            }
            stk[sp++] = z;
          } else {
            const float2 x = stk[--sp];
            float2 z;
            switch (ins.op) {
              case MXR_EW_CONJ: z = make_float2(x.x, -x.y); break;
              case MXR_EW_ABS2: z = make_float2(x.x * x.x + x.y * x.y, 0.f); break;
              case MXR_EW_ABS: z = make_float2(sqrtf(x.x * x.x + x.y * x.y), 0.f); break;
              case MXR_EW_SQRT: z = make_float2(sqrtf(x.x), 0.f); break;
              case MXR_EW_NEG: z = make_float2(-x.x, -x.y); break;
              case MXR_EW_REAL: z = make_float2(x.x, 0.f); break;
              case MXR_EW_IMAG: z = make_float2(x.y, 0.f); break;
              case MXR_EW_EXP: {
                const float m = expf(x.x);
                z = make_float2(m * cosf(x.y), m * sinf(x.y));
                break;
              }
              case MXR_EW_LOG: z = make_float2(logf(sqrtf(x.x * x.x + x.y * x.y)), atan2f(x.y, x.x)); break;
              case MXR_EW_COS: z = make_float2(cosf(x.x), 0.f); break;
              default: z = make_float2(sinf(x.x), 0.f); break;
            }
            stk[sp++] = z;
          }
        }
      }
    }
    const float2 v = stk[0];
    const int64_t o = ooff + e * a.ost[r - 1];
    switch (a.odtype) {
      case MXR_C64: reinterpret_cast<float2 *>(a.out)[o] = v; break;
      case MXR_F32: reinterpret_cast<float *>(a.out)[o] = v.x; break;
      default: reinterpret_cast<int *>(a.out)[o] = (int)v.x; break;
    }
  }
}

static mxr_status ew_validate_program(const mxr_ew_instr *prog, int nprog, int nin, int nconst,
                                      int rank) {
  MXR_REQUIRE(nprog >= 1 && nprog <= EW_MAXPROG, MXR_INVALID_PARAMETER,
              "ew: program length %d not in [1, %d]", nprog, EW_MAXPROG);
  int sp = 0;
  for (int pc = 0; pc < nprog; ++pc) {
    const mxr_ew_instr &i = prog[pc];
    int pops = 0, pushes = 1;
    switch (i.op) {
      case MXR_EW_LOAD:
        MXR_REQUIRE(i.a >= 0 && i.a < nin, MXR_INVALID_PARAMETER, "ew: LOAD of input %d", i.a);
        break;
      case MXR_EW_CONST:
        MXR_REQUIRE(i.a >= 0 && i.a < nconst, MXR_INVALID_PARAMETER, "ew: CONST %d", i.a);
        break;
      case MXR_EW_HAMMING:
      case MXR_EW_INDEX:
        MXR_REQUIRE(i.a >= 0 && i.a < rank, MXR_INVALID_PARAMETER, "ew: dim %d", i.a);
        break;
      case MXR_EW_SELECT: pops = 3; break;
      default:
        if (i.op >= MXR_EW_ADD && i.op <= MXR_EW_MIN) pops = 2;
        else if (i.op >= MXR_EW_CONJ && i.op <= MXR_EW_SIN) pops = 1;
        else {
          set_error(MXR_INVALID_PARAMETER, "ew: unknown opcode %d", i.op);
          return MXR_INVALID_PARAMETER;
        }
    }
    MXR_REQUIRE(sp >= pops, MXR_INVALID_PARAMETER, "ew: stack underflow at instruction %d", pc);
    sp += pushes - pops;
    MXR_REQUIRE(sp <= EW_STACK, MXR_INVALID_PARAMETER, "ew: stack deeper than %d", EW_STACK);
  }
  MXR_REQUIRE(sp == 1, MXR_INVALID_PARAMETER, "ew: program leaves %d values", sp);
  return MXR_SUCCESS;
}

// Opcodes whose operands must be real: they read only the real part
// (ordering, pow, max/min, sqrt, cos, sin, the selector of SELECT).  A value
// is complex when it comes from a complex input or constant, or from an
// arithmetic / conj / neg / exp / log of a complex value.  Rejecting complex
// operands here keeps e.g. sqrt(complex tensor) from silently returning the
// real-part result.
static mxr_status ew_check_types(const mxr_ew_instr *prog, int nprog, const int *idt,
                                 const float *consts) {
  bool cplx[EW_STACK + 3];
  int sp = 0;
  auto need_real = [&](bool c, int pc) -> mxr_status {
    MXR_REQUIRE(!c, MXR_INVALID_TYPE,
                "ew: opcode %d at instruction %d takes a real operand, got a complex one",
                prog[pc].op, pc);
    return MXR_SUCCESS;
  };
  mxr_status st;
  for (int pc = 0; pc < nprog; ++pc) {
    const mxr_ew_instr &i = prog[pc];
    switch (i.op) {
      case MXR_EW_LOAD: cplx[sp++] = idt[i.a] == MXR_C64; break;
      case MXR_EW_CONST: cplx[sp++] = consts[2 * i.a + 1] != 0.f; break;
      case MXR_EW_HAMMING:
      case MXR_EW_INDEX: cplx[sp++] = false; break;
      case MXR_EW_SELECT: {
        const bool c = cplx[--sp], y = cplx[--sp], x = cplx[--sp];
        if ((st = need_real(c, pc))) return st;
        cplx[sp++] = x || y;
        break;
      }
      default:
        if (i.op >= MXR_EW_ADD && i.op <= MXR_EW_MIN) {
          const bool y = cplx[--sp], x = cplx[--sp];
          const bool arith = i.op <= MXR_EW_DIV, eq = i.op == MXR_EW_EQ;
          if (!arith && !eq && (st = need_real(x || y, pc))) return st;
          cplx[sp++] = arith && (x || y);
        } else {
          const bool x = cplx[--sp];
          switch (i.op) {
            case MXR_EW_SQRT:
            case MXR_EW_COS:
            case MXR_EW_SIN:
              if ((st = need_real(x, pc))) return st;
              cplx[sp++] = false;
              break;
            case MXR_EW_ABS2:
            case MXR_EW_ABS:
            case MXR_EW_REAL:
            case MXR_EW_IMAG: cplx[sp++] = false; break;
            default: cplx[sp++] = x; break;  // CONJ, NEG, EXP, LOG
          }
        }
    }
  }
  return MXR_SUCCESS;
}

// ------------------------------------------------------ expression compiler
// The default path generates one HIP kernel per expression shape (program,
// operand dtypes, rank, which constants are real) as straight-line source,
// compiles it with hipRTC for the current device and caches the module: the
// MatX analogue is the compile-time fused kernel of kernel.h and the NVRTC
// string generator of executors/jit_cuda.h:325-420.  Values known to be real
// at generation time (real inputs, real constants, comparisons, |.|^2, ...)
// stay single floats, so a real expression does real arithmetic only.
// Constants, pointers, shape and strides are kernel arguments, so a new
// constant value or view reuses the cached kernel.  The interpreter above
// runs when the compiler is unavailable or MXR_EW_JIT=0.
struct EwjArgs {
  void *out;
  const void *in[EW_MAXIN];
  long long shape[MXR_MAX_RANK];
  long long ost[MXR_MAX_RANK];
  long long ist[EW_MAXIN][MXR_MAX_RANK];
  long long cblocks;  // workgroups per row (EWJ_COLS columns each)
  long long pad;
  float cst[2 * EW_MAXCONST];
};
// Columns of one row per generated-kernel workgroup: 256 lanes x 4.
constexpr int EWJ_COLS = 1024;

// Device source shared by every generated kernel: the argument block (the
// same layout as EwjArgs) and the complex helpers of the interpreter.
static const char *kEwjPrelude = R"(
struct EwjArgs {
  void *out;
  const void *in[EWJ_MAXIN];
  long long shape[EWJ_RANKMAX];
  long long ost[EWJ_RANKMAX];
  long long ist[EWJ_MAXIN][EWJ_RANKMAX];
  long long cblocks;
  long long pad;
  float cst[2 * EWJ_MAXCONST];
};
__device__ __forceinline__ float2 mk(float x, float y) { float2 r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {
  const float d = b.x * b.x + b.y * b.y;
  return mk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ float ham(long long i, int n) {
  if (n <= 1) return 1.0f;
  return 0.54f - 0.46f * cosf(6.28318530717958647692f * (float)i / (float)(n - 1));
}
// x^y for a base x that is a constant of the expression, with lx = log2(x)
// in double formed once per kernel: exp2 of t = y lx, t split into float
// high and low parts (exp2(th + tl) = exp2(th) (1 + tl ln 2) to second
// order, |tl| <= 2^-17), so the result carries the rounding of one exp2f and
// one multiply.  Bases that are not positive and finite, x = 1 and products
// outside the normal float exponent range take powf.
__device__ __forceinline__ float pow_cl(float x, float y, double lx) {
  const double t = (double)y * lx;
  if (!(x > 0.f) || x == 1.f || !__builtin_isfinite(x) || !(t > -126.0 && t < 126.0)) return powf(x, y);
  const float th = (float)t;
  const float tl = (float)(t - (double)th);
  return exp2f(th) * fmaf(tl, 0.693147180559945309f, 1.0f);
}
)";

namespace {
struct EwjVal {
  std::string name;
  bool real;
  int cst = -1;  // the constant it is, when it is one (a CONST operand)
};

std::string ewj_c(const EwjVal &v) { return v.real ? "mk(" + v.name + ", 0.f)" : v.name; }
std::string ewj_re(const EwjVal &v) { return v.real ? v.name : v.name + ".x"; }

// Straight-line source of one program; `creal` marks real constants.
std::string ewj_source(const mxr_ew_instr *prog, int nprog, const int *idt, int nin, int odt,
                       int rank, unsigned creal) {
  std::ostringstream k;
  k << "#define EWJ_MAXIN " << EW_MAXIN << "\n#define EWJ_RANKMAX " << MXR_MAX_RANK
    << "\n#define EWJ_MAXCONST " << EW_MAXCONST << "\n"
    << kEwjPrelude
    // a workgroup owns EWJ_COLS columns of one row: the row's coordinates and
    // base offsets are workgroup-uniform (scalar unit), and consecutive lanes
    // take consecutive columns (coalesced)
    << "extern \"C\" __global__ __launch_bounds__(256) void mxr_ew_jit(EwjArgs a) {\n"
    << "  const long long row = (long long)blockIdx.x / a.cblocks;\n"
    << "  const long long col0 = ((long long)blockIdx.x - row * a.cblocks) * " << EWJ_COLS << ";\n"
    << "  long long idx[" << rank << "];\n"
    << "  long long rem = row;\n"
    << "  idx[" << rank - 1 << "] = 0;\n";
  for (int d = rank - 2; d >= 0; --d)
    k << "  idx[" << d << "] = rem % a.shape[" << d << "]; rem /= a.shape[" << d << "];\n";
  k << "  long long oo = 0;\n";
  for (int i = 0; i < nin; ++i) k << "  long long io" << i << " = 0;\n";
  for (int d = 0; d < rank - 1; ++d) {
    k << "  oo += idx[" << d << "] * a.ost[" << d << "];\n";
    for (int i = 0; i < nin; ++i)
      k << "  io" << i << " += idx[" << d << "] * a.ist[" << i << "][" << d << "];\n";
  }
  // The EWJ_COLS / 256 columns of a thread are unrolled in the source:
  // every operand load of every column is issued first (a column past the
  // row's end reads the row's last column instead), then the arithmetic,
  // then the stores of the columns inside the row.  Loads written after a
  // per-column exit would wait out one memory latency per column.
  constexpr int NE = EWJ_COLS / 256;
  k << "  const long long ncol = a.shape[" << rank - 1 << "];\n";
  for (int e = 0; e < NE; ++e)
    k << "  const long long c" << e << "r = col0 + threadIdx.x + " << e * 256 << ";\n"
      << "  const bool ok" << e << " = c" << e << "r < ncol;\n"
      << "  const long long c" << e << " = ok" << e << " ? c" << e << "r : ncol - 1;\n";
  auto load_name = [](int e, int pc) { return "l" + std::to_string(e) + "_" + std::to_string(pc); };
  for (int e = 0; e < NE; ++e)
    for (int pc = 0; pc < nprog; ++pc) {
      if (prog[pc].op != MXR_EW_LOAD) continue;
      const int i = prog[pc].a;
      const std::string off = "io" + std::to_string(i) + " + c" + std::to_string(e) + " * a.ist[" +
                              std::to_string(i) + "][" + std::to_string(rank - 1) + "]";
      const std::string p = "a.in[" + std::to_string(i) + "]", n = load_name(e, pc);
      if (idt[i] == MXR_C64) k << "  const float2 " << n << " = ((const float2 *)" << p << ")[" << off << "];\n";
      else if (idt[i] == MXR_F32) k << "  const float " << n << " = ((const float *)" << p << ")[" << off << "];\n";
      else k << "  const float " << n << " = (float)((const int *)" << p << ")[" << off << "];\n";
    }
  std::set<int> pow_logs;  // constants whose log2 the kernel has formed
  for (int e = 0; e < NE; ++e) {
    const std::string c = "c" + std::to_string(e);
    std::vector<EwjVal> st;
    int nv = 0;
    auto def = [&](bool real, const std::string &expr) {
      EwjVal v{"v" + std::to_string(e) + "_" + std::to_string(nv++), real};
      k << "  const " << (real ? "float " : "float2 ") << v.name << " = " << expr << ";\n";
      st.push_back(v);
    };
    auto pos = [&](int d) {
      return d == rank - 1 ? c : "idx[" + std::to_string(d) + "]";
    };
    for (int pc = 0; pc < nprog; ++pc) {
      const mxr_ew_instr &ins = prog[pc];
      switch (ins.op) {
        case MXR_EW_LOAD: {
          const int i = ins.a;
          st.push_back(EwjVal{load_name(e, pc), idt[i] != MXR_C64});
          break;
        }
        case MXR_EW_CONST: {
          const std::string c = std::to_string(2 * ins.a);
          if (creal >> ins.a & 1u) def(true, "a.cst[" + c + "]");
          else def(false, "mk(a.cst[" + c + "], a.cst[" + c + " + 1])");
          st.back().cst = ins.a;
          break;
        }
        case MXR_EW_HAMMING: def(true, "ham(" + pos(ins.a) + ", " + std::to_string(ins.b) + ")"); break;
        case MXR_EW_INDEX: def(true, "(float)" + pos(ins.a)); break;
        case MXR_EW_SELECT: {
          const EwjVal c = st.back(); st.pop_back();
          const EwjVal y = st.back(); st.pop_back();
          const EwjVal x = st.back(); st.pop_back();
          if (x.real && y.real) def(true, "(" + ewj_re(c) + " != 0.f) ? " + x.name + " : " + y.name);
          else def(false, "(" + ewj_re(c) + " != 0.f) ? " + ewj_c(x) + " : " + ewj_c(y));
          break;
        }
        default: {
          if (ins.op >= MXR_EW_ADD && ins.op <= MXR_EW_MIN) {
            const EwjVal y = st.back(); st.pop_back();
            const EwjVal x = st.back(); st.pop_back();
            const bool rr = x.real && y.real;
            const std::string X = ewj_c(x), Y = ewj_c(y), xr = ewj_re(x), yr = ewj_re(y);
            switch (ins.op) {
              case MXR_EW_ADD: rr ? def(true, x.name + " + " + y.name) : def(false, "cadd(" + X + ", " + Y + ")"); break;
              case MXR_EW_SUB: rr ? def(true, x.name + " - " + y.name) : def(false, "csub(" + X + ", " + Y + ")"); break;
              case MXR_EW_MUL:
                if (rr) def(true, x.name + " * " + y.name);
                else if (x.real) def(false, "mk(" + x.name + " * " + y.name + ".x, " + x.name + " * " + y.name + ".y)");
                else if (y.real) def(false, "mk(" + x.name + ".x * " + y.name + ", " + x.name + ".y * " + y.name + ")");
                else def(false, "cmul(" + x.name + ", " + y.name + ")");
                break;
              case MXR_EW_DIV:
                if (rr) def(true, x.name + " / " + y.name);
                else if (y.real) def(false, "mk(" + X + ".x / " + y.name + ", " + X + ".y / " + y.name + ")");
                else def(false, "(" + y.name + ".y == 0.f) ? mk(" + X + ".x / " + y.name + ".x, " + X + ".y / " +
                                     y.name + ".x) : cdiv(" + X + ", " + y.name + ")");
                break;
              case MXR_EW_POW:
                if (x.real && x.cst >= 0) {
                  // log2 of the constant base once per kernel (before its
                  // first use, in scope for every later column)
                  const std::string lg = "plg" + std::to_string(x.cst);
                  if (pow_logs.insert(x.cst).second)
                    k << "  const double " << lg << " = a.cst[" << 2 * x.cst << "] > 0.f ? log2((double)a.cst["
                      << 2 * x.cst << "]) : 0.0;\n";
                  def(true, "pow_cl(" + xr + ", " + yr + ", " + lg + ")");
                } else {
                  def(true, "powf(" + xr + ", " + yr + ")");
                }
                break;
              case MXR_EW_GT: def(true, "(" + xr + " > " + yr + ") ? 1.f : 0.f"); break;
              case MXR_EW_LT: def(true, "(" + xr + " < " + yr + ") ? 1.f : 0.f"); break;
              case MXR_EW_GE: def(true, "(" + xr + " >= " + yr + ") ? 1.f : 0.f"); break;
              case MXR_EW_LE: def(true, "(" + xr + " <= " + yr + ") ? 1.f : 0.f"); break;
              case MXR_EW_EQ:
                if (rr) def(true, "(" + x.name + " == " + y.name + ") ? 1.f : 0.f");
                else def(true, "(" + X + ".x == " + Y + ".x && " + X + ".y == " + Y + ".y) ? 1.f : 0.f");
                break;
              case MXR_EW_MAX: def(true, "fmaxf(" + xr + ", " + yr + ")"); break;
              default: def(true, "fminf(" + xr + ", " + yr + ")"); break;
            }
          } else {
            const EwjVal x = st.back(); st.pop_back();
            const std::string X = ewj_c(x), xr = ewj_re(x), n = x.name;
            switch (ins.op) {
              case MXR_EW_CONJ: x.real ? def(true, n) : def(false, "mk(" + n + ".x, -" + n + ".y)"); break;
              case MXR_EW_ABS2: x.real ? def(true, n + " * " + n) : def(true, n + ".x * " + n + ".x + " + n + ".y * " + n + ".y"); break;
              case MXR_EW_ABS:
                x.real ? def(true, "sqrtf(" + n + " * " + n + ")")
                       : def(true, "sqrtf(" + n + ".x * " + n + ".x + " + n + ".y * " + n + ".y)");
                break;
              case MXR_EW_SQRT: def(true, "sqrtf(" + xr + ")"); break;
              case MXR_EW_NEG: x.real ? def(true, "-" + n) : def(false, "mk(-" + n + ".x, -" + n + ".y)"); break;
              case MXR_EW_REAL: def(true, xr); break;
              case MXR_EW_IMAG: x.real ? def(true, "0.f") : def(true, n + ".y"); break;
              case MXR_EW_EXP:
                if (x.real) def(true, "expf(" + n + ")");
                else def(false, "mk(expf(" + n + ".x) * cosf(" + n + ".y), expf(" + n + ".x) * sinf(" + n + ".y))");
                break;
              case MXR_EW_LOG:
                def(false, "mk(logf(sqrtf(" + X + ".x * " + X + ".x + " + X + ".y * " + X + ".y)), atan2f(" + X +
                               ".y, " + X + ".x))");
                break;
              case MXR_EW_COS: def(true, "cosf(" + xr + ")"); break;
              default: def(true, "sinf(" + xr + ")"); break;
            }
          }
        }
      }
    }
    const EwjVal v = st.back();
    const std::string o = "oo + " + c + " * a.ost[" + std::to_string(rank - 1) + "]";
    k << "  if (ok" << e << ") ";
    if (odt == MXR_C64) k << "((float2 *)a.out)[" << o << "] = " << ewj_c(v) << ";\n";
    else if (odt == MXR_F32) k << "((float *)a.out)[" << o << "] = " << ewj_re(v) << ";\n";
    else k << "((int *)a.out)[" << o << "] = (int)" << ewj_re(v) << ";\n";
  }
  k << "}\n";
  return k.str();
}

struct EwjKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  bool failed = false;
};
std::mutex g_ewj_mu;
std::map<std::pair<int, std::string>, EwjKernel> g_ewj;

bool ewj_enabled() {
  static const bool on = [] {
    const char *e = getenv("MXR_EW_JIT");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Compiles `src` for the current device with the library's code-generation
// flags (packed FP32 instructions off).  Empty on failure.
std::vector<char> ewj_compile(const std::string &src) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id()) != hipSuccess) return {};
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  hiprtcProgram p;
  if (hiprtcCreateProgram(&p, src.c_str(), "mxr_ew_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return {};
  const char *opts[] = {arch.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  std::vector<char> code;
  if (hiprtcCompileProgram(p, 6, opts) == HIPRTC_SUCCESS) {
    size_t n = 0;
    if (hiprtcGetCodeSize(p, &n) == HIPRTC_SUCCESS && n > 0) {
      code.resize(n);
      if (hiprtcGetCode(p, code.data()) != HIPRTC_SUCCESS) code.clear();
    }
  }
  (void)hiprtcDestroyProgram(&p);
  return code;
}
}  // namespace

// The generated kernel of this expression shape, compiled on first use;
// nullptr when the compiler is off or failed (the interpreter runs then).
static hipFunction_t ewj_kernel(const mxr_ew_instr *prog, int nprog, const int *idt, int nin,
                                int odt, int rank, unsigned creal) {
  if (!ewj_enabled()) return nullptr;
  std::ostringstream key;
  key << rank << '|' << odt << '|' << creal << '|';
  for (int i = 0; i < nin; ++i) key << idt[i] << ',';
  key << '|';
  for (int i = 0; i < nprog; ++i) key << prog[i].op << ':' << prog[i].a << ':' << prog[i].b << ';';
  std::lock_guard<std::mutex> g(g_ewj_mu);
  EwjKernel &kk = g_ewj[{device_id(), key.str()}];
  if (kk.fn || kk.failed) return kk.fn;
  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<char> code = ewj_compile(ewj_source(prog, nprog, idt, nin, odt, rank, creal));
  if (code.empty() || hipModuleLoadData(&kk.mod, code.data()) != hipSuccess ||
      hipModuleGetFunction(&kk.fn, kk.mod, "mxr_ew_jit") != hipSuccess) {
    kk.fn = nullptr;
    kk.failed = true;
    log_msg(LOG_WARN, "ew: generated kernel failed to build; the interpreter runs [%s]", key.str().c_str());
  } else {
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    log_msg(LOG_DEBUG, "ew: compiled a generated kernel in %.1f ms [%s]", ms, key.str().c_str());
  }
  return kk.fn;
}

static void clear_ewj() {
  std::lock_guard<std::mutex> g(g_ewj_mu);
  for (auto &kv : g_ewj) {
    if (!kv.second.mod) continue;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first.first);
    (void)hipModuleUnload(kv.second.mod);
    (void)hipSetDevice(cur);
  }
  g_ewj.clear();
}
static CacheClearRegistrar g_reg_ewj(clear_ewj);

mxr_status ew_impl(const mxr_tensor *out, const mxr_tensor *const *ins, int nin,
                   const mxr_ew_instr *prog, int nprog, const float *consts, int nconst,
                   hipStream_t stream) {
  mxr_status st;
  if ((st = check_tensor(out, "ew output", 0, 1, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(nin >= 0 && nin <= EW_MAXIN, MXR_INVALID_PARAMETER, "ew: %d inputs", nin);
  MXR_REQUIRE(nconst >= 0 && nconst <= EW_MAXCONST, MXR_INVALID_PARAMETER, "ew: %d constants",
              nconst);
  MXR_REQUIRE(prog != nullptr, MXR_INVALID_PARAMETER, "ew: null program");
  if ((st = ew_validate_program(prog, nprog, nin, nconst, out->rank))) return st;
  EwArgs a{};
  a.out = out->data;
  a.odtype = out->dtype;
  a.rank = out->rank;
  for (int d = 0; d < out->rank; ++d) {
    a.shape[d] = out->shape[d];
    a.ost[d] = out->stride[d];
  }
  a.nin = nin;
  for (int k = 0; k < nin; ++k) {
    if ((st = check_tensor(ins[k], "ew input", 0, out->rank, out->rank))) return st;
    for (int d = 0; d < out->rank; ++d)
      MXR_REQUIRE(ins[k]->shape[d] == out->shape[d] || ins[k]->stride[d] == 0, MXR_INVALID_SIZE,
                  "ew: input %d dim %d extent %lld does not match the output %lld", k, d,
                  (long long)ins[k]->shape[d], (long long)out->shape[d]);
    a.in[k].p = ins[k]->data;
    a.in[k].dtype = ins[k]->dtype;
    for (int d = 0; d < out->rank; ++d) a.in[k].st[d] = ins[k]->stride[d];
  }
  {
    int dts[EW_MAXIN] = {0};
    for (int k = 0; k < nin; ++k) dts[k] = ins[k]->dtype;
    if ((st = ew_check_types(prog, nprog, dts, consts))) return st;
  }
  a.nprog = nprog;
  for (int i = 0; i < nprog; ++i) a.prog[i] = prog[i];
  for (int i = 0; i < nconst; ++i) a.cst[i] = make_float2(consts[2 * i], consts[2 * i + 1]);
  const int64_t n_last = out->shape[out->rank - 1];
  if (numel(*out) == 0) return MXR_SUCCESS;
  a.runs_per_row = (n_last + EW_EPT - 1) / EW_EPT;
  a.total_runs = (numel(*out) / n_last) * a.runs_per_row;
  const int64_t blocks = (a.total_runs + 255) / 256;
  MXR_REQUIRE(blocks <= 0x7fffffff, MXR_INVALID_SIZE, "ew: too large");
  int idt[EW_MAXIN] = {0};
  for (int k = 0; k < nin; ++k) idt[k] = ins[k]->dtype;
  unsigned creal = 0;
  for (int i = 0; i < nconst; ++i)
    if (consts[2 * i + 1] == 0.f) creal |= 1u << i;
  if (hipFunction_t fn = ewj_kernel(prog, nprog, idt, nin, out->dtype, out->rank, creal)) {
    EwjArgs j{};
    j.out = out->data;
    for (int k = 0; k < nin; ++k) {
      j.in[k] = ins[k]->data;
      for (int d = 0; d < out->rank; ++d) j.ist[k][d] = ins[k]->stride[d];
    }
    for (int d = 0; d < out->rank; ++d) {
      j.shape[d] = out->shape[d];
      j.ost[d] = out->stride[d];
    }
    j.cblocks = (n_last + EWJ_COLS - 1) / EWJ_COLS;
    const int64_t jgrid = (numel(*out) / n_last) * j.cblocks;
    MXR_REQUIRE(jgrid <= 0x7fffffff, MXR_INVALID_SIZE, "ew: too large");
    for (int i = 0; i < 2 * nconst; ++i) j.cst[i] = consts[i];
    void *params[] = {&j};
    ProfScope ps("ew_jit_kernel", stream);
    MXR_HIP_TRY(hipModuleLaunchKernel(fn, (unsigned)jgrid, 1, 1, 256, 1, 1, 0, stream, params, nullptr));
    return MXR_SUCCESS;
  }
  {
    ProfScope ps("ew_kernel", stream);
    ew_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// ------------------------------------------------------------------ sum
// Two launches: each of up to SUM_BLOCKS workgroups reduces a grid-strided
// share of the input in double into one partial (float2), then one
// workgroup adds the partials in order.  The view's shape and strides
// travel in the argument block.
constexpr int SUM_BLOCKS = 512;

struct SumArgs {
  const void *in;
  int is_c, rank;
  int64_t n;
  int64_t shape[MXR_MAX_RANK], stride[MXR_MAX_RANK];
  float2 *partial;  // SUM_BLOCKS entries
  void *out;
  int out_c;
  int64_t blocks;   // partials written by stage 0
};

__global__ __launch_bounds__(256) void sum_kernel(SumArgs a, int stage) {
  __shared__ double sr[256], si[256];
  double ar = 0.0, ai = 0.0;
  if (stage == 0) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < a.n; e += (int64_t)gridDim.x * 256) {
      int64_t rem = e, off = 0;
      for (int d = a.rank - 1; d >= 0; --d) {
        const int64_t s = a.shape[d];
        off += (rem % s) * a.stride[d];
        rem /= s;
      }
      if (a.is_c) {
        const float2 z = reinterpret_cast<const float2 *>(a.in)[off];
        ar += z.x;
        ai += z.y;
      } else {
        ar += reinterpret_cast<const float *>(a.in)[off];
      }
    }
  } else {
    for (int e = threadIdx.x; e < (int)a.blocks; e += 256) {
      ar += a.partial[e].x;
      ai += a.partial[e].y;
    }
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (stage == 0) {
      a.partial[blockIdx.x] = make_float2((float)sr[0], (float)si[0]);
    } else if (a.out_c) {
      *reinterpret_cast<float2 *>(a.out) = make_float2((float)sr[0], (float)si[0]);
    } else {
      *reinterpret_cast<float *>(a.out) = (float)sr[0];
    }
  }
}

// Validates a sum call and fills everything but the partials buffer.
static mxr_status sum_args(const mxr_tensor *out, const mxr_tensor *in, SumArgs *a) {
  mxr_status st;
  if ((st = check_tensor(in, "sum input", 0, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(out, "sum output", 0, 0, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(in->dtype == MXR_C64 || in->dtype == MXR_F32, MXR_INVALID_TYPE,
              "sum: input must be c64 or f32");
  MXR_REQUIRE(out->dtype == in->dtype, MXR_INVALID_TYPE, "sum: output dtype must match input");
  MXR_REQUIRE(numel(*out) == 1, MXR_INVALID_SIZE, "sum: output must hold one element");
  *a = SumArgs{};
  a->in = in->data;
  a->is_c = in->dtype == MXR_C64;
  a->rank = in->rank;
  a->n = numel(*in);
  for (int d = 0; d < in->rank; ++d) {
    a->shape[d] = in->shape[d];
    a->stride[d] = in->stride[d];
  }
  a->out = out->data;
  a->out_c = out->dtype == MXR_C64;
  int64_t blocks = (a->n + 255) / 256;
  if (blocks > SUM_BLOCKS) blocks = SUM_BLOCKS;
  if (blocks < 1) blocks = 1;
  a->blocks = blocks;
  return MXR_SUCCESS;
}

static mxr_status sum_launch(const SumArgs &a, hipStream_t stream) {
  ProfScope ps("sum_kernel", stream);
  sum_kernel<<<dim3((unsigned)a.blocks), dim3(256), 0, stream>>>(a, 0);
  sum_kernel<<<dim3(1), dim3(256), 0, stream>>>(a, 1);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

mxr_status sum_impl(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream) {
  SumArgs a;
  mxr_status st = sum_args(out, in, &a);
  if (st) return st;
  a.partial = (float2 *)stream_workspace("sum", sizeof(float2) * SUM_BLOCKS, stream);
  if (!a.partial) return workspace_failure("sum");
  return sum_launch(a, stream);
}

// ----------------------------------------------- sum without allocation
// mxr_sum_noalloc keeps its partials in this file-scope array: one copy per
// GPU, resident until the process ends, shared by every call on that GPU.
// Stage 0 writes slots [0, blocks) and stage 1 reads only those, so nothing
// has to start at zero.
static __device__ __align__(256) float2 scratch[SUM_BLOCKS];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float2 *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};  // per device: the address of `scratch`
}  // namespace

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address).
static mxr_status scratch_address(float2 **p) {
  const int dev = device_id();
  MXR_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, MXR_NOT_SUPPORTED,
              "sum_noalloc: device %d beyond the %d devices served", dev, SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *q = nullptr;
    MXR_HIP_TRY(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<float2 *>(q);
  }
  *p = g_scratch_ptr[dev];
  return MXR_SUCCESS;
}

mxr_status sum_noalloc_impl(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream) {
  SumArgs a;
  mxr_status st = sum_args(out, in, &a);
  if (st) return st;
  // the first launch's grid is capped at the scratch's SUM_BLOCKS slots, so
  // every input size fits; the check guards that invariant
  MXR_REQUIRE(a.blocks >= 1 && a.blocks <= SUM_BLOCKS, MXR_INVALID_SIZE,
              "sum_noalloc: %lld partials exceed the %d-slot scratch", (long long)a.blocks,
              SUM_BLOCKS);
  if ((st = scratch_address(&a.partial))) return st;
  return sum_launch(a, stream);
}

}  // namespace mxr

extern "C" {

mxr_status mxr_ew_eval(const mxr_tensor *out, const mxr_tensor *const *ins, int nin,
                       const mxr_ew_instr *prog, int nprog, const float *consts, int nconst,
                       hipStream_t stream) {
  return mxr::guarded("mxr_ew_eval", [&] {
    return mxr::ew_impl(out, ins, nin, prog, nprog, consts, nconst, stream);
  });
}

mxr_status mxr_sum(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream) {
  return mxr::guarded("mxr_sum", [&] { return mxr::sum_impl(out, in, stream); });
}

mxr_status mxr_sum_noalloc(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream) {
  return mxr::guarded("mxr_sum_noalloc", [&] { return mxr::sum_noalloc_impl(out, in, stream); });
}

}  // extern "C"
