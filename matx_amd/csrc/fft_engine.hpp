// Workgroup FFT engine for gfx950: radix-2/4/8/16 Stockham autosort passes
// with butterflies held in VGPRs and the data exchanged through LDS between
// passes.
//
// Stockham DIT pass with radix R over a length-N sequence, with sub-transform
// length Ls done so far (Ls = product of earlier radices):
//   butterfly j in [0, N/R), k = j mod Ls
//   inputs  a_m = X[j + m*N/R]                     m = 0..R-1
//   a_m *= W_{Ls*R}^{m*k}                           (W_n = e^{-2 pi i/n} forward)
//   b = DFT_R(a)
//   Y[(j-k)*R + q*Ls + k] = b_q                     q = 0..R-1
// After the last pass Y is in natural order.  Twiddles come from tables built
// on the host in double precision and rounded once to float (one table per
// (Ls, R) pass, k-major with stride R so a thread reads its R entries as one
// contiguous run).
#pragma once

#include "mxr_internal.hpp"

namespace mxr {

// ------------------------------------------------- in-register small DFTs
// Forward W_n^e = exp(-2 pi i e / n); inverse uses the conjugate.
template <bool INV>
__device__ __forceinline__ float2 mul_w4(float2 a) {  // * W_4^1 = -i (fwd) / +i (inv)
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2 &a0, float2 &a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_w4<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// Multiply by W_16^e (compile-time e in [0,16)), trivial factors special-cased.
template <int E, bool INV>
__device__ __forceinline__ float2 mul_w16(float2 a) {
  constexpr int e = E & 15;
  constexpr float C1 = 0.92387953251128675613f;  // cos(pi/8)
  constexpr float S1 = 0.38268343236508977173f;  // sin(pi/8)
  constexpr float H = 0.70710678118654752440f;   // sqrt(1/2)
  if constexpr (e == 0) return a;
  if constexpr (e == 4) return mul_w4<INV>(a);
  if constexpr (e == 8) return make_float2(-a.x, -a.y);
  if constexpr (e == 12) return mul_w4<!INV>(a);
  // general: angle = -2 pi e/16 (fwd); cos/sin by symmetry from pi/8 multiples
  constexpr float cs[16] = {1.f, C1, H, S1, 0.f, -S1, -H, -C1, -1.f, -C1, -H, -S1, 0.f, S1, H, C1};
  constexpr float sn[16] = {0.f, S1, H, C1, 1.f, C1, H, S1, 0.f, -S1, -H, -C1, -1.f, -C1, -H, -S1};
  // W = cos(2pi e/16) - i sin(2pi e/16) forward
  constexpr float wr = cs[e];
  constexpr float wi_f = -sn[e];
  const float wi = INV ? -wi_f : wi_f;
  if constexpr (e == 2 || e == 6 || e == 10 || e == 14) {
    // (x + iy)(wr + i wi) with |wr| = |wi| = H
    float sr = a.x * wr - a.y * wi;
    float si = a.x * wi + a.y * wr;
    return make_float2(sr, si);
  }
  return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2 (&a)[8]) {
  // 8 = 2 x 4: a[m1 + 2 m2]; inner DFT4 over m2, twiddle W8^{m1 c}, DFT2 over m1.
  dft4<INV>(a[0], a[2], a[4], a[6]);
  dft4<INV>(a[1], a[3], a[5], a[7]);
  a[3] = mul_w16<2, INV>(a[3]);
  a[5] = mul_w16<4, INV>(a[5]);
  a[7] = mul_w16<6, INV>(a[7]);
  float2 y[8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float2 u = a[2 * c], v = a[2 * c + 1];
    y[c] = cadd(u, v);
    y[c + 4] = csub(u, v);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = y[i];
}

// a * W_16^e for e in {2, 6} equals H * hrot(a), H = sqrt(1/2), with hrot
// made of additions only (W_16^2 = H (1 - i), W_16^6 = H (-1 - i) forward;
// the inverse conjugates).
template <int E, bool INV>
__device__ __forceinline__ float2 hrot(float2 a) {
  static_assert(E == 2 || E == 6, "hrot: e in {2, 6}");
  if constexpr (E == 2)
    return INV ? make_float2(a.x - a.y, a.x + a.y) : make_float2(a.x + a.y, a.y - a.x);
  else
    return INV ? make_float2(-a.x - a.y, a.x - a.y) : make_float2(a.y - a.x, -a.x - a.y);
}
// p = u + H r, m = u - H r, as fused multiply-adds
__device__ __forceinline__ void hpair(float2 u, float2 r, float2 &p, float2 &m) {
  constexpr float H = 0.70710678118654752440f;
  p = make_float2(fmaf(H, r.x, u.x), fmaf(H, r.y, u.y));
  m = make_float2(fmaf(-H, r.x, u.x), fmaf(-H, r.y, u.y));
}

template <bool INV>
__device__ __forceinline__ void dft16(float2 (&a)[16]) {
  // 16 = 4 x 4: inner DFT4 over m2 (stride 4), twiddle W16^{m1 c}, DFT4 over
  // m1.  Z[m1][c] sits in a[m1 + 4c]; column c's outputs go to y[c + 4r].
  // The W16^2 / W16^6 twiddles are folded into the column butterflies as
  // H-scaled fused multiply-adds.
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1) dft4<INV>(a[m1], a[m1 + 4], a[m1 + 8], a[m1 + 12]);
  float2 y[16];
  {  // c = 0: no twiddles
    float2 z0 = a[0], z1 = a[1], z2 = a[2], z3 = a[3];
    dft4<INV>(z0, z1, z2, z3);
    y[0] = z0;
    y[4] = z1;
    y[8] = z2;
    y[12] = z3;
  }
  {  // c = 1: W^1, W^2, W^3
    const float2 z1 = mul_w16<1, INV>(a[5]), z3 = mul_w16<3, INV>(a[7]);
    float2 t0, t1;
    hpair(a[4], hrot<2, INV>(a[6]), t0, t1);
    const float2 t2 = cadd(z1, z3), t3 = mul_w4<INV>(csub(z1, z3));
    y[1] = cadd(t0, t2);
    y[9] = csub(t0, t2);
    y[5] = cadd(t1, t3);
    y[13] = csub(t1, t3);
  }
  {  // c = 2: W^2, W^4, W^6 = W^2 W^4
    const float2 z2 = mul_w4<INV>(a[10]);
    const float2 t0 = cadd(a[8], z2), t1 = csub(a[8], z2);
    const float2 q = mul_w4<INV>(a[11]);
    const float2 ru = hrot<2, INV>(cadd(a[9], q)), rv = hrot<2, INV>(csub(a[9], q));
    hpair(t0, ru, y[2], y[10]);
    hpair(t1, mul_w4<INV>(rv), y[6], y[14]);
  }
  {  // c = 3: W^3, W^6, W^9
    const float2 z1 = mul_w16<3, INV>(a[13]), z3 = mul_w16<9, INV>(a[15]);
    float2 t0, t1;
    hpair(a[12], hrot<6, INV>(a[14]), t0, t1);
    const float2 t2 = cadd(z1, z3), t3 = mul_w4<INV>(csub(z1, z3));
    y[3] = cadd(t0, t2);
    y[11] = csub(t0, t2);
    y[7] = cadd(t1, t3);
    y[15] = csub(t1, t3);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = y[i];
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&a)[R]) {
  if constexpr (R == 2) dft2<INV>(a[0], a[1]);
  else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 8) dft8<INV>(a);
  else if constexpr (R == 16) dft16<INV>(a);
  else static_assert(R == 2 || R == 4 || R == 8 || R == 16, "radix");
}

// --------------------------------------------------------------- plans
// Radix sequence per power-of-two size (forward order).  The inverse of a
// fused forward->inverse pair runs the reversed sequence, so its first pass
// has the forward's last radix and can reuse the registers directly.
struct RadixSeq {
  int n;
  int r[6];
};
__host__ __device__ constexpr RadixSeq radix_seq(int N) {
  switch (N) {
    case 1: return {0, {0}};
    case 2: return {1, {2}};
    case 4: return {1, {4}};
    case 8: return {1, {8}};
    case 16: return {1, {16}};
    case 32: return {2, {8, 4}};
    case 64: return {2, {16, 4}};
    case 128: return {2, {16, 8}};
    case 256: return {2, {16, 16}};
    case 512: return {3, {16, 8, 4}};
    case 1024: return {3, {16, 16, 4}};
    case 2048: return {3, {16, 16, 8}};
    case 4096: return {3, {16, 16, 16}};
    case 8192: return {4, {16, 16, 8, 4}};
    case 16384: return {4, {16, 16, 16, 4}};
    default: return {-1, {0}};
  }
}
__host__ __device__ constexpr RadixSeq reversed(RadixSeq s) {
  RadixSeq o{s.n, {0}};
  for (int i = 0; i < s.n; ++i) o.r[i] = s.r[s.n - 1 - i];
  return o;
}
// Offset (in float2) of pass p's twiddle table inside a sequence's table.
__host__ __device__ constexpr int64_t tw_offset(RadixSeq s, int p) {
  int64_t off = 0, Ls = 1;
  for (int i = 0; i < p; ++i) {
    off += (Ls > 1) ? Ls * s.r[i] : 0;
    Ls *= s.r[i];
  }
  return off;
}
__host__ __device__ constexpr int64_t tw_size(RadixSeq s) { return tw_offset(s, s.n); }
__host__ __device__ constexpr int pass_ls(RadixSeq s, int p) {
  int Ls = 1;
  for (int i = 0; i < p; ++i) Ls *= s.r[i];
  return Ls;
}

// LDS padding: one spare float2 every 16, which makes the stride-R writes of
// the early passes bank-conflict free for R in {4, 8, 16}.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_padded_len(int n) { return n + (n >> 4) + 1; }

// --------------------------------------------------------- pass helpers
template <int R, bool INV>
__device__ __forceinline__ void apply_twiddles(float2 (&a)[R], const float2 *__restrict__ tw, int k) {
  // tw points at the pass table; entries k*R + m, m = 1..R-1 (entry 0 holds
  // 1 and is skipped).  Two twiddles per 16-byte load.
  const float4 *t4 = reinterpret_cast<const float4 *>(tw + (int64_t)k * R);
#pragma unroll
  for (int h = 0; h < R / 2; ++h) {
    float4 w = t4[h];
    if (h > 0) a[2 * h] = cmul(a[2 * h], make_float2(w.x, INV ? -w.y : w.y));
    a[2 * h + 1] = cmul(a[2 * h + 1], make_float2(w.z, INV ? -w.w : w.w));
  }
}

// Host: build the twiddle table of a radix sequence for length N in double
// precision, rounded once to float.  Layout per pass p (Ls > 1):
// entry k*R + m = exp(-2 pi i m k / (Ls R)).
void build_twiddles(int N, RadixSeq s, float2 *out);

// Per-device cache of twiddle tables keyed by (N, reversed?).  Returns a
// device pointer owned by the library.
const float2 *twiddles_device(int N, bool reversed_seq);

// v[m] *= W^{mk}, m = 1..15, from one power-base row (W^k, W^2k, W^4k,
// W^8k): the other powers are products of at most four of them (at most
// three extra roundings).  inv conjugates (inverse transform).
__device__ __forceinline__ void tw_from_base(float2 (&v)[16], float4 b0, float4 b1, bool inv) {
  float2 w1 = make_float2(b0.x, b0.y), w2 = make_float2(b0.z, b0.w);
  float2 w4 = make_float2(b1.x, b1.y), w8 = make_float2(b1.z, b1.w);
  if (inv) {
    w1.y = -w1.y;
    w2.y = -w2.y;
    w4.y = -w4.y;
    w8.y = -w8.y;
  }
  float2 w[16];
  w[1] = w1;
  w[2] = w2;
  w[4] = w4;
  w[8] = w8;
  w[3] = cmul(w2, w1);
  w[5] = cmul(w4, w1);
  w[6] = cmul(w4, w2);
  w[7] = cmul(w[6], w1);
  w[9] = cmul(w8, w1);
  w[10] = cmul(w8, w2);
  w[11] = cmul(w[10], w1);
  w[12] = cmul(w8, w4);
  w[13] = cmul(w[12], w1);
  w[14] = cmul(w[12], w2);
  w[15] = cmul(w[14], w1);
#pragma unroll
  for (int m = 1; m < 16; ++m) v[m] = cmul(v[m], w[m]);
}

// Per-device cache of radix-16 power-base rows for the pass with sub-length
// Ls of an N-point transform: row k (k < Ls) holds W^k, W^2k, W^4k, W^8k with
// W = e^{-2 pi i / (16 Ls)}, each from double precision rounded once.
const float4 *pow_base_device(int N, int Ls);

// Raw buffer access (MUBUF) for the block kernels: a resource covering
// `bytes` bytes from `base`.  The hardware range check returns zeros for
// loads at offsets at or past `bytes` and drops such stores, so a block that
// runs past the end of its row needs no per-element tests.  Offsets are
// byte offsets in the vector operand (the scalar offset stays 0: the range
// check does not cover it).
// The base and size must be uniform across the workgroup (they are read from
// the first lane, so the resource lives in scalar registers).
using buf_rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ buf_rsrc make_rsrc(const void *base, int bytes) {
  const uint64_t b = reinterpret_cast<uint64_t>(base);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  void *ub = reinterpret_cast<void *>(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(ub, (short)0, __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ __forceinline__ float2 buf_load2(buf_rsrc r, int voff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}
__device__ __forceinline__ float4 buf_load4(buf_rsrc r, int voff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}
__device__ __forceinline__ void buf_store2(buf_rsrc r, int voff, float2 v) {
  using u2 = __attribute__((ext_vector_type(2))) unsigned;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, 0, 0);
}

// Workgroup barrier that is also a compiler fence: no load or store is moved
// across it, so twiddle and spectrum loads stay inside their pass instead of
// being hoisted (and held in VGPRs) across the whole kernel.
__device__ __forceinline__ void wg_barrier() {
  asm volatile("" ::: "memory");
  __syncthreads();
  asm volatile("" ::: "memory");
}

// ------------------------------------------------------------ layouts
// ROW: TB transforms stored one after another, each padded (lds_pad);
//      consecutive lanes take consecutive butterflies of one transform.
// COL: element i of transform (column) t at i*TB + t; consecutive lanes take
//      consecutive columns, so every LDS access is unit stride across lanes.
template <int N, int TB, bool COL>
struct Layout {
  static constexpr int TSTRIDE = lds_padded_len(N);
  static constexpr int LDS_ELEMS = COL ? N * TB : TB * TSTRIDE;
  __device__ static __forceinline__ int addr(int t, int i) {
    if constexpr (COL) return i * TB + t;
    else return t * TSTRIDE + lds_pad(i);
  }
  template <int R>
  __device__ static __forceinline__ void bfly(int g, int &t, int &j) {
    constexpr int NB = N / R;
    if constexpr (COL && TB % 64 == 0) {
      // a wave spans whole rows of columns: j is the same on every lane
      // (kept in a scalar register, so twiddle reads become scalar loads)
      t = g % TB;
      j = __builtin_amdgcn_readfirstlane(g / TB);
    } else if constexpr (COL) {
      t = g % TB;
      j = g / TB;
    } else {
      t = g / NB;
      j = g % NB;
    }
  }
};

template <int N, bool REV>
struct Seq {
  static constexpr RadixSeq s = REV ? reversed(radix_seq(N)) : radix_seq(N);
  static constexpr int n = s.n;
  template <int P>
  static constexpr int R = s.r[P];
  template <int P>
  static constexpr int LS = pass_ls(s, P);
  template <int P>
  static constexpr int64_t TWOFF = tw_offset(s, P);
};

// One Stockham pass: butterflies of this thread held in v[u][m].
template <class LY, int R, int N, int NT, int TB>
struct Pass {
  static constexpr int NB = N / R;
  static constexpr int TOT = NB * TB;
  static constexpr int BPT = (TOT + NT - 1) / NT;
  float2 v[BPT][R];

  __device__ __forceinline__ static bool active(int u) {
    return (TOT % NT == 0) || (int)(threadIdx.x + u * NT) < TOT;
  }
  // load(t, i) -> float2 : element i of transform t
  template <class LoadF>
  __device__ __forceinline__ void load(LoadF &&f) {
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
      if (!active(u)) continue;
      int t, j;
      LY::template bfly<R>(threadIdx.x + u * NT, t, j);
#pragma unroll
      for (int m = 0; m < R; ++m) v[u][m] = f(t, j + m * NB);
    }
  }
  // twiddle + DFT_R in registers
  template <int LS, bool INV>
  __device__ __forceinline__ void compute(const float2 *__restrict__ tw) {
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
      if (!active(u)) continue;
      if constexpr (LS > 1) {
        int t, j;
        LY::template bfly<R>(threadIdx.x + u * NT, t, j);
        apply_twiddles<R, INV>(v[u], tw, j & (LS - 1));
      }
      dft<R, INV>(v[u]);
    }
  }
  // f(u, q, t, i) for the output positions store<LS> visits, value
  // v[u][q] going to element i of transform t (to fetch per-output data
  // ahead of the store, or to store with it)
  template <int LS, class F>
  __device__ __forceinline__ static void outputs(F &&f) {
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
      if (!active(u)) continue;
      int t, j;
      LY::template bfly<R>(threadIdx.x + u * NT, t, j);
      const int k = j & (LS - 1);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) f(u, q, t, base + q * LS);
    }
  }
  // store(t, i, value) for the Stockham output positions
  template <int LS, class StoreF>
  __device__ __forceinline__ void store(StoreF &&f) {
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
      if (!active(u)) continue;
      int t, j;
      LY::template bfly<R>(threadIdx.x + u * NT, t, j);
      const int k = j & (LS - 1);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) f(t, base + q * LS, v[u][q]);
    }
  }
};

// Runs passes [P, PE) of sequence SQ with both input and output in LDS.
// Pre: LDS holds pass P's input, visible to all threads (barrier passed).
// Post: LDS holds pass PE-1's output, visible to all threads.
template <class LY, class SQ, int N, int NT, int TB, bool INV, int P, int PE>
__device__ __forceinline__ void lds_passes(float2 *lds, const float2 *__restrict__ tw) {
  if constexpr (P < PE) {
    constexpr int R = SQ::template R<P>;
    constexpr int LS = SQ::template LS<P>;
    Pass<LY, R, N, NT, TB> ps;
    ps.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();
    ps.template compute<LS, INV>(tw + SQ::template TWOFF<P>);
    ps.template store<LS>([&](int t, int i, float2 x) { lds[LY::addr(t, i)] = x; });
    wg_barrier();
    lds_passes<LY, SQ, N, NT, TB, INV, P + 1, PE>(lds, tw);
  }
}

}  // namespace mxr
