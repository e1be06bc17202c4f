/*
 * mxr.h -- C ABI of libmxr.so, the MI355X (gfx950) radar signal-processing
 * library behind the MatX-compatible operator front end (include/matx_amd.h).
 *
 * Conventions
 *   - All tensor memory is caller-owned device memory (hipMalloc / managed).
 *     The library never frees a caller pointer.  It owns only opaque plans,
 *     twiddle tables and waveform spectra, released by mxr_clear_caches().
 *   - Every call is asynchronous on the caller's hipStream_t; nothing
 *     synchronises the device.  Plans are created (and their device buffers
 *     allocated) on first use, so warm a shape once before hipGraph capture;
 *     a call that would have to grow a per-stream workspace while its stream
 *     is being captured fails with MXR_NOT_SUPPORTED instead of allocating.
 *   - Every entry returns an mxr_status (0 = success).  On failure the
 *     thread-local text from mxr_last_error() says what was wrong.  Shape and
 *     size errors are detected on the host before any launch, as MatX throws
 *     matxException before launching (include/matx/core/error.h:55-74).
 *   - Re-entrant from several host threads; plan lookup is mutex-guarded
 *     (cf. include/matx/core/cache.h:221-257).
 *
 * Citations are to the MatX sources (paths relative to the MatX repository
 * root) whose behaviour each entry point replaces.
 */
#ifndef MXR_H_
#define MXR_H_

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MXR_MAX_RANK 8

/* Status codes mirror matxError_t (include/matx/core/error.h:55-74). */
typedef enum mxr_status {
  MXR_SUCCESS = 0,
  MXR_NOT_SUPPORTED = 1,
  MXR_INVALID_PARAMETER = 2,
  MXR_INVALID_DIM = 3,
  MXR_INVALID_SIZE = 4,
  MXR_OUT_OF_MEMORY = 5,
  MXR_INVALID_TYPE = 6,
  MXR_HIP_ERROR = 7,
  MXR_ALIAS_ERROR = 8
} mxr_status;

typedef enum mxr_dtype {
  MXR_C64 = 1, /* complex<float>, interleaved re,im */
  MXR_F32 = 2,
  MXR_I32 = 3
} mxr_dtype;

/* A strided view: element (i0,...,i{rank-1}) lives at
 * data + sum_d i_d*stride[d] elements (cf. tensor_impl.h:91-160). */
typedef struct mxr_tensor {
  void *data;
  int32_t dtype;
  int32_t rank;
  int64_t shape[MXR_MAX_RANK];
  int64_t stride[MXR_MAX_RANK];
} mxr_tensor;

/* Convolution modes (include/matx/core/operator_options.h:52-60). */
typedef enum mxr_conv_mode { MXR_CONV_FULL = 0, MXR_CONV_SAME = 1, MXR_CONV_VALID = 2 } mxr_conv_mode;

/* FFT normalisation (FFTNorm, include/matx/core/operator_options.h:45-49). */
typedef enum mxr_fft_norm { MXR_NORM_BACKWARD = 0, MXR_NORM_FORWARD = 1, MXR_NORM_ORTHO = 2 } mxr_fft_norm;
typedef enum mxr_fft_dir { MXR_FFT_FORWARD = 0, MXR_FFT_INVERSE = 1 } mxr_fft_dir;

/* ------------------------------------------------------------------ misc */
const char *mxr_last_error(void);
int mxr_version(void);
/* Frees every cached plan, twiddle table, per-stream waveform spectrum and
 * per-stream workspace (cf. ClearCachesAndAllocations, test/main.cu).  It
 * synchronises the device first.  Caller-created plans (mxr_pc_plan_*,
 * mxr_ragged_plan_*) stay valid: while any of them is alive the twiddle
 * tables they point at are kept.  A hipGraph captured before the call may
 * reference per-stream workspaces and spectra that it frees: re-capture
 * after clearing.  Workspaces that had to grow keep their old buffers until
 * this call, so earlier captures stay valid until then. */
mxr_status mxr_clear_caches(void);

/* Kernel timing, the analogue of the executor timers start_timer /
 * stop_timer / get_time_ms (include/matx/executors/cuda_executor_common.h:
 * 142-169).  While enabled, every kernel the library launches is bracketed by
 * a pair of hipEvents recorded on the stream it is launched on.
 * mxr_profile_query sums the recorded launches of kernels whose name contains
 * `name` (waiting for their events); mxr_profile_reset drops the records. */
mxr_status mxr_profile_enable(int on);
mxr_status mxr_profile_reset(void);
mxr_status mxr_profile_query(const char *name, int64_t *count, double *total_ms);

/* Library log (MatX's logger, include/matx/core/log.h:110-401: MATX_LOG_LEVEL
 * / MATX_LOG_DEST).  Levels TRACE 0, DEBUG 1, INFO 2, WARN 3, ERROR 4,
 * FATAL 5, OFF 6; start-up values from MXR_LOG_LEVEL (a name or a number)
 * and MXR_LOG_DEST ("stderr", "stdout" or a file path, appended).  Default:
 * OFF, stderr (MatX: ERROR, stdout) -- failures are returned as statuses
 * already, and stdout belongs to the caller.  When enabled, every failed
 * call logs its message at ERROR, plan and cache events at DEBUG and every
 * kernel launch at TRACE.  mxr_log_write adds a caller message. */
mxr_status mxr_log_set_level(int level);
int mxr_log_get_level(void);
mxr_status mxr_log_set_dest(const char *dest);
void mxr_log_write(int level, const char *message);

/* Trace ranges (MatX's NVTX ranges, include/matx/core/nvtx.h:48-168, emitted
 * here through roctx for rocprofv3 --marker-trace).  Levels are MatX's
 * matx_nvxtLogLevels; a range is emitted when its level is at most the trace
 * level (default API, as nvtx.h:93; MXR_TRACE_LEVEL overrides it: a number
 * or NONE / USER / API / INTERNAL / ALL).  The library opens an INTERNAL
 * range around every kernel launch, named after the kernel.  roctx is loaded
 * at the first range that passes the filter; without it ranges are no-ops
 * (mxr_trace_available says which).
 * mxr_trace_push / mxr_trace_pop: a nested range (MATX_NVTX_START); push
 * returns 1 when it opened one, which the caller then pops.
 * mxr_trace_range_start / _end: a range ended by id from anywhere
 * (MATX_NVTX_START_RANGE / MATX_NVTX_END_RANGE, nvtx.h:122-134); *id < 0
 * asks for a fresh id, written back.  Starting an id that is open is an
 * error; ending an id that is not open does nothing. */
enum { MXR_TRACE_NONE = 0, MXR_TRACE_USER = 1, MXR_TRACE_API = 2, MXR_TRACE_INTERNAL = 3, MXR_TRACE_ALL = 0xFFFFFF };
mxr_status mxr_trace_set_level(int level);
int mxr_trace_get_level(void);
int mxr_trace_available(void);
int mxr_trace_push(const char *name, int level);
void mxr_trace_pop(void);
mxr_status mxr_trace_range_start(const char *name, int level, int *id);
mxr_status mxr_trace_range_end(int id);

/* DLPack interop (MatX: tensor_t::ToDlPack / make_tensor(DLManagedTensor*),
 * include/matx/core/tensor.h:1471-1555, make_tensor.h:1136-1190).
 * mxr_dl_tensor has the memory layout of DLPack's DLTensor (data, device
 * {type, id}, ndim, dtype {code, bits, lanes}, shape, strides, byte_offset),
 * so a producer's DLTensor* (or the first member of its DLManagedTensor) can
 * be passed as is.  mxr_tensor_from_dlpack fills a descriptor of the same
 * memory, without copying: ROCm device memory (device type 10) only; complex64
 * (code 5, 64 bits), float32 (2, 32) or int32 (0, 32), one lane; NULL strides
 * mean compact row-major; rank 0..8.  Ownership stays with the producer. */
typedef struct mxr_dl_tensor {
  void *data;
  int32_t device_type, device_id;
  int32_t ndim;
  uint8_t dtype_code, dtype_bits;
  uint16_t dtype_lanes;
  int64_t *shape;
  int64_t *strides;
  uint64_t byte_offset;
} mxr_dl_tensor;
mxr_status mxr_tensor_from_dlpack(mxr_tensor *out, const mxr_dl_tensor *src);

/* DLPack's DLManagedTensor layout: the tensor, the producer's context and
 * the deleter the consumer calls once when it is done. */
typedef struct mxr_dl_managed {
  mxr_dl_tensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(struct mxr_dl_managed *self);
} mxr_dl_managed;

/* Export (tensor_t::ToDlPack, include/matx/core/tensor.h:1471-1555): a
 * DLManagedTensor describing t's memory, no copy, with device {device_type,
 * device_id} (10 = ROCm device memory, 11 = ROCm pinned host memory) and
 * t's shape and element strides.  The library allocates the struct and its
 * shape / stride arrays; the consumer's single call of its deleter frees
 * them and then calls release(release_ctx) (release may be NULL), where the
 * producer drops the reference it took for the export, as MatX's export
 * holds a reference to the tensor's storage.  complex64, float32, int32. */
mxr_status mxr_tensor_to_dlpack(const mxr_tensor *t, int32_t device_type, int32_t device_id,
                                void (*release)(void *ctx), void *release_ctx,
                                mxr_dl_managed **out);

/* ------------------------------------------------------------------ FFT */
/* Batched 1-D complex-to-complex FFT along `axis`.
 * n = transform size; 0 means in->shape[axis].  The input is zero-padded or
 * truncated to n, and out->shape[axis] must equal n; every other dimension of
 * in and out must match (batch dims).  Any strides; in == out allowed.
 * Power-of-two n from 1 to 16384: one workgroup per transform (Stockham
 * radix-16/8/4/2 passes in LDS); from 32768 to 2^24: two passes of short
 * transforms through a per-stream work matrix of batch * n elements (the
 * four-step factorisation).  Any other n from 3 to 8192 through Bluestein's
 * chirp-z identity (one power-of-two convolution of length >= 2n - 1 in
 * LDS), and from 8193 to 2^23 through Bluestein with two-pass convolution
 * transforms (two work matrices).  The work matrices grow on first use:
 * run a size once before capturing it in a graph.  (cuFFT's arbitrary sizes,
 * which MatX inherits, are mixed-radix.)
 * Replaces fft_impl / ifft_impl (include/matx/transforms/fft/fft_cuda.h:680-774,
 * size rules include/matx/operators/fft.h:111-199 and fft_common.h:39-140,
 * normalisation fft_cuda.h:126-188). */
mxr_status mxr_fft_c2c(const mxr_tensor *out, const mxr_tensor *in, int64_t n,
                       int dir, int norm, int axis, hipStream_t stream);

/* Real-to-complex FFT along `axis` (rfft, include/matx/operators/fft.h:580-626):
 * in f32, out c64 with n/2 + 1 bins along the axis; n = 0 means
 * in->shape[axis]; the input is zero-padded or truncated to n.  Sizes, norms
 * and strides as mxr_fft_c2c (forward direction). */
mxr_status mxr_fft_r2c(const mxr_tensor *out, const mxr_tensor *in, int64_t n,
                       int norm, int axis, hipStream_t stream);

/* Complex-to-real inverse FFT along `axis` (irfft, include/matx/operators/
 * fft.h:686-740): bins 0..n/2 of `in` (zero-padded or truncated) extended
 * as X[n-k] = conj(X[k]); out f32 with n samples along the axis; n = 0
 * means 2 * (in->shape[axis] - 1).  Sizes, norms and strides as
 * mxr_fft_c2c (inverse direction). */
mxr_status mxr_fft_c2r(const mxr_tensor *out, const mxr_tensor *in, int64_t n,
                       int norm, int axis, hipStream_t stream);

/* fft_conv: out = ifft(fft(x, n) .* conj(fft(w, n)))  (conj_w = 1, circular
 * cross-correlation, the benchmarked operator) or ifft(fft(x,n) .* fft(w,n))
 * (conj_w = 0, circular convolution), along the last axis of x.
 * x: [..., S] c64; w: [L] c64 (one waveform for every row) or [..., L] with
 * x's batch dims (one waveform per row, n in [1024, 8192]); out: [..., nout]
 * c64 with nout <= n, holding lags 0..nout-1.  There is no MatX operator of this name; it is the idiom of
 * examples/fft_conv.cu:90-108 and the conj-multiply of
 * examples/simple_radar_pipeline.h:295-300. */
mxr_status mxr_fft_conv(const mxr_tensor *out, const mxr_tensor *x,
                        const mxr_tensor *w, int64_t n, int conj_w,
                        hipStream_t stream);

/* Adjoint of mxr_fft_conv with one waveform per row (the backward pass of
 * a loss through fft_conv; no MatX counterpart: MatX has no autodiff).
 * gy: [..., nout] c64, the gradient of a real loss w.r.t. y in the
 * convention g = dL/dRe(y) + i dL/dIm(y).  x: [..., S], w: [..., L] as in the
 * forward call.  Writes, for the rows' correlation (conj_w = 1)
 *   gx = IFFT_n(FFT_n(gy) .* FFT_n(w))[0..S),  gw = IFFT_n(FFT_n(x) .* conj(FFT_n(gy)))[0..L)
 * and for convolution (conj_w = 0)
 *   gx = IFFT_n(FFT_n(gy) .* conj(FFT_n(w))),  gw = IFFT_n(FFT_n(gy) .* conj(FFT_n(x))).
 * gx or gw may be NULL.  n in [1024, 8192]. */
mxr_status mxr_fft_conv_adjoint(const mxr_tensor *gx, const mxr_tensor *gw, const mxr_tensor *gy,
                                const mxr_tensor *x, const mxr_tensor *w, int64_t n, int conj_w,
                                hipStream_t stream);

/* Gradient w.r.t. s of r = fft_conv(s, s, n) (each row's matched filter to
 * itself, correlation form), i.e. gx + gw of mxr_fft_conv_adjoint with
 * x = w = s, in one pass: gs = IFFT_n(FFT_n(s) .* 2 Re(FFT_n(gy)))[0..S).
 * gy: [..., nout] c64; s, gs: [..., S] c64; n in [1024, 8192]. */
mxr_status mxr_autocorr_adjoint(const mxr_tensor *gs, const mxr_tensor *gy, const mxr_tensor *s,
                                int64_t n, hipStream_t stream);

/* Pulse compression stage, examples/simple_radar_pipeline.h:273-301:
 *   w1 = w .* hamming(L) (if MXR_PC_WINDOW); w2 = w1/||w1|| (if
 *   MXR_PC_NORMALIZE); y[..., r] = (1/n) ifft(fft(x, n) .* conj(fft(w2, n)))[r]
 *   for r < y->shape[last].
 * x: [..., S] c64 (last axis unit stride); y: [..., nout] c64, nout <= n.
 * y->data == x->data is allowed (in place, as MatX transforms inputView in
 * place); then each row is finished before any of it is overwritten.
 * MXR_PC_WRITEBACK stores w2 back into w, reproducing MatX's in-place
 * re-windowing of waveformView on every call (:288-294). */
#define MXR_PC_WINDOW 1
#define MXR_PC_NORMALIZE 2
#define MXR_PC_WRITEBACK 4
mxr_status mxr_pulse_compress(const mxr_tensor *y, const mxr_tensor *x,
                              const mxr_tensor *w, int64_t n, int flags,
                              hipStream_t stream);

/* Pulse-compression plan: the waveform preparation of PulseCompression
 * (examples/simple_radar_pipeline.h:276-296: window, unit-energy scale,
 * conj(fft)) done once per waveform instead of once per cube, as a library
 * FFT plan is made once per shape (cf. the cuFFT plan cache,
 * include/matx/transforms/fft/fft_cuda.h:285-291).
 * create: rows of S samples (x's last axis), waveforms of L samples,
 *   transform size n (power of two), nout lags kept, flags MXR_PC_WINDOW |
 *   MXR_PC_NORMALIZE.  Picks the block size and allocates the spectrum on
 *   the current device.
 * set_waveform: computes the spectrum of w ([L] c64) on `stream`; later
 *   executes on other streams must be ordered after it by the caller.
 * execute: y[..., r] = mxr_pulse_compress's result for r < nout, from the
 *   stored spectrum.  x: [..., S] c64, y: [..., nout] c64, no overlap.
 * destroy: frees the plan (synchronising the device). */
typedef struct mxr_pc_plan_s *mxr_pc_plan;
mxr_status mxr_pc_plan_create(mxr_pc_plan *plan, int64_t S, int64_t L, int64_t n, int64_t nout,
                              int flags);
mxr_status mxr_pc_plan_set_waveform(mxr_pc_plan plan, const mxr_tensor *w, hipStream_t stream);
mxr_status mxr_pc_plan_execute(mxr_pc_plan plan, const mxr_tensor *y, const mxr_tensor *x,
                               hipStream_t stream);
mxr_status mxr_pc_plan_destroy(mxr_pc_plan plan);

/* Gradient of mxr_pc_plan_execute with respect to the raw waveform, summed
 * over every row (no MatX counterpart: the backward of PulseCompression,
 * examples/simple_radar_pipeline.h:276-300, for waveform design).  gy:
 * [..., nout] c64 = dL/dRe y + i dL/dIm y for the rows of x ([..., S]); w:
 * [L] the waveform the forward used; gw: [L] c64 receives dL/dw through the
 * plan's window and normalisation.  With w2 the filter the forward applied,
 *   g2[l] = sum_rows sum_{m < nout} x[(m + l) mod n] conj(gy[m]),
 *   w1 = h .* w, nu = ||w1||, gw = h .* (g2/nu - w1 Re<g2, w1>/nu^3)
 * (h = 1 without MXR_PC_WINDOW; gw = h .* g2 without MXR_PC_NORMALIZE).
 * Runs over the plan's overlap-save blocks, recomputing the spectra of x
 * and gy; the per-workgroup partial sums are reduced in a fixed order, so
 * the result is deterministic. */
mxr_status mxr_pc_plan_waveform_grad(mxr_pc_plan plan, const mxr_tensor *gw, const mxr_tensor *gy,
                                     const mxr_tensor *x, const mxr_tensor *w, hipStream_t stream);

/* ------------------------------------------------------------ pulse axis */
/* tpc_mode value for mxr_doppler: y already holds the canceller output; all
 * of its rows feed the FFT (MatX's DopplerProcessing stage on tpcView,
 * simple_radar_pipeline.h:344-357, window on rows < win_rows). */
#define MXR_TPC_NONE 3

/* Three-pulse canceller + Doppler window + Doppler FFT (+ |.|^2), fused.
 * y: [C, P, R] c64 (R axis unit stride).  tpc_mode MXR_CONV_SAME reproduces
 * MatX's stage dataflow (simple_radar_pipeline.h:321-357: SAME [1,-2,1] over
 * all P pulses, Hamming(win_rows) on rows < win_rows, the remaining rows
 * unwindowed, zero rows up to nd, nd-point FFT along pulses);
 * MXR_CONV_VALID the numpy model (01_radar.py:65-78: rows 1..P-2).
 * rd_out: [C, nd, R] c64 range-Doppler map (may be NULL); it may be exactly
 * y (same data, shape and strides: the in-place stage form), never a partial
 * overlap.  pow_out: [C, nd, R] f32 |rd|^2 (may be NULL).  At least one
 * output is required. */
mxr_status mxr_doppler(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                       const mxr_tensor *y, int tpc_mode, int64_t win_rows,
                       int64_t nd, hipStream_t stream);

/* Doppler stage with an arbitrary real window (the differentiable form):
 * rd = FFT_nd(taps .* x) along pulses, pow = |rd|^2.  x: [C, P, R] c64 (no
 * canceller), taps: [P] f32; outputs as mxr_doppler. */
mxr_status mxr_doppler_taps(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                            const mxr_tensor *x, const mxr_tensor *taps, int64_t nd,
                            hipStream_t stream);

/* Adjoint of mxr_doppler_taps' power output (no MatX counterpart: MatX has
 * no autodiff).  gpow: [C, nd, R] f32 = dL/dpow.  With D = FFT_nd(taps.*x),
 * gu = F^H (2 gpow .* D) (unnormalised inverse DFT, rows p < P):
 *   gx = taps .* gu  ([C, P, R] c64, g = dL/dRe + i dL/dIm)
 *   gtaps[p] = sum_{c,r} Re(conj(gu[c,p,r]) x[c,p,r])  ([P] f32)
 * D is recomputed from x (not stored).  Either output may be NULL. */
/* Adjoint of mxr_doppler's power output with respect to y (no MatX
 * counterpart; the gradient of ThreePulseCanceller + DopplerProcessing +
 * abs2, examples/simple_radar_pipeline.h:321-357,398).  Same tpc_mode,
 * win_rows and nd as the forward call.  gpow: [C, nd, R] f32 = dL/dpow (any
 * strides, 0 included); gy: [C, P, R] c64 = dL/dRe y + i dL/dIm y.  With
 * t = window .* canceller(y), D = FFT_nd(t), s = window .* F^H(2 gpow .* D):
 *   gy[p] = s[p-o+1] - 2 s[p-o] + s[p-o-1]   (o = 1 for VALID, else 0)
 * or gy = s without a canceller.  D is recomputed from y (checkpointing:
 * no spectrum is stored between forward and backward).  Every tile is
 * transformed (mxr_doppler_taps_adjoint, for sparse design losses, skips
 * tiles whose gpow is all zero). */
mxr_status mxr_doppler_adjoint(const mxr_tensor *gy, const mxr_tensor *gpow, const mxr_tensor *y,
                               int tpc_mode, int64_t win_rows, int64_t nd, hipStream_t stream);

mxr_status mxr_doppler_taps_adjoint(const mxr_tensor *gtaps, const mxr_tensor *gx,
                                    const mxr_tensor *gpow, const mxr_tensor *x,
                                    const mxr_tensor *taps, int64_t nd, hipStream_t stream);

/* Sidelobe loss heads for waveform/window design, forward and adjoint (no
 * MatX counterpart).  All reductions in double.
 * mxr_isl: r [B, n] c64 (unit row stride) -> isl [B] f32,
 *   isl_b = (sum_l |r_b[l]|^2 - |r_b[0]|^2) / |r_b[0]|^2.
 * mxr_isl_adjoint: gr [B, n] c64 = d(sum_b gisl_b isl_b)/dr.
 * mxr_psl: pw [C, nd, R] f32, targets [T, 3] i32 (channel, range bin,
 *   Doppler bin kd) -> psl [T] f32,
 *   psl_t = (sum_{k: circular |k - kd| > ml} pw[c, k, r]^q)^(1/q) / pw[c, kd, r]
 *   (NaN for a target outside pw).
 * mxr_psl_adjoint: gpow [C, nd, R] f32 (contiguous) = d(sum_t gpsl_t psl_t)/dpw;
 *   zero outside the target columns, columns shared by targets summed in
 *   target order. */
mxr_status mxr_isl(const mxr_tensor *isl, const mxr_tensor *r, hipStream_t stream);
/* The phase-only waveform's ISL in one kernel each way (no MatX counterpart):
 * mxr_phase_isl: phi [B, L] f32 -> isl [B] f32, the ISL of the n-point
 *   autocorrelation fft_conv(s, s, n) of s = exp(j phi) (1 <= L <= n, n a
 *   power of two in [1024, 8192]); equal to mxr_isl of that autocorrelation.
 * mxr_phase_isl_adjoint: gphi [B, L] f32 = d(sum_b gisl_b isl_b)/dphi. */
mxr_status mxr_phase_isl(const mxr_tensor *isl, const mxr_tensor *phi, int64_t n, hipStream_t stream);
mxr_status mxr_phase_isl_adjoint(const mxr_tensor *gphi, const mxr_tensor *gisl, const mxr_tensor *phi,
                                 int64_t n, hipStream_t stream);
mxr_status mxr_isl_adjoint(const mxr_tensor *gr, const mxr_tensor *gisl, const mxr_tensor *r,
                           hipStream_t stream);
mxr_status mxr_psl(const mxr_tensor *psl, const mxr_tensor *pw, const mxr_tensor *targets, int ml,
                   int q, hipStream_t stream);
mxr_status mxr_psl_adjoint(const mxr_tensor *gpow, const mxr_tensor *gpsl, const mxr_tensor *pw,
                           const mxr_tensor *targets, int ml, int q, hipStream_t stream);
/* The taps gradient of sum_t gpsl_t psl_t for pw = |FFT_nd(taps .* x)|^2
 * along pulses (mxr_doppler_taps' power), taken from the target columns
 * alone (no MatX counterpart): gtaps [P] f32 from gpsl [T], the saved pw
 * [C, nd, R], x [C, P, R] c64 and taps [P]; equal to mxr_psl_adjoint
 * followed by mxr_doppler_taps_adjoint's gtaps, without the map-sized
 * gradient.  Shares per target are summed in target order. */
mxr_status mxr_psl_taps_adjoint(const mxr_tensor *gtaps, const mxr_tensor *gpsl, const mxr_tensor *pw,
                                const mxr_tensor *x, const mxr_tensor *taps, const mxr_tensor *targets,
                                int ml, int q, hipStream_t stream);

/* --------------------------------------------------------- ragged batch */
/* Many independent pulse-compression + Doppler jobs of uneven shapes in a
 * few persistent launches (one per block-size class and stage), for capture
 * into one hipGraph.  No MatX counterpart: MatX runs such jobs as separate
 * transform calls (fft_impl, include/matx/transforms/fft/fft_cuda.h:680-774).
 * Per job:
 *   y  = pulse compression of x ([C, P, S] c64, unit sample stride) with
 *        the window-normalised waveform w ([L] c64, unit stride) at
 *        n = next_pow2(S), lags r < R = S - L + 1 (as mxr_pulse_compress
 *        with MXR_PC_WINDOW | MXR_PC_NORMALIZE), kept in plan scratch;
 *   pw = |FFT_nd(hamming_P .* y)|^2 along pulses, nd = next_pow2(P)
 *        ([C, nd, R] f32, unit range stride), as mxr_doppler(pw, NULL, y,
 *        MXR_TPC_NONE, P, nd).
 * Strides are in elements.  513 <= S <= 16384 (block sizes 1024 to 16384),
 * 2 <= P <= 1024. */
typedef struct mxr_ragged_job {
  const void *x;
  int64_t C, P, S;
  int64_t x_cs, x_ps; /* channel and pulse strides of x */
  const void *w;
  int64_t L;
  void *pw;
  int64_t pw_cs, pw_ks; /* channel and Doppler-bin strides of pw */
} mxr_ragged_job;
typedef struct mxr_ragged_plan_s *mxr_ragged_plan;
/* Pick each size class's overlap-save block size by timing the candidates
 * on the device (cached per process); otherwise a flop-count cost model. */
#define MXR_RAGGED_AUTOTUNE 1
/* Run each class as a fixed grid of resident workgroups looping over the
 * class's items (a persistent launch) instead of one workgroup per item. */
#define MXR_RAGGED_PERSISTENT 2
/* Copies the job table, allocates the scratch, runs the autotuner on
 * `stream` (synchronously). */
mxr_status mxr_ragged_plan_create(mxr_ragged_plan *plan, const mxr_ragged_job *jobs, int njobs,
                                  int flags, hipStream_t stream);
/* Enqueues every job: waveform spectra, pulse compression, Doppler.  The
 * jobs of one Doppler size form a chain (their block-size classes' spectrum
 * and pulse-compression launches, then their Doppler launch); chains run
 * concurrently on the plan's own helper streams, forked from and joined back
 * into `stream` with events (parallel branches when captured), so one
 * chain's Doppler pass overlaps another's pulse compression.  No allocation
 * and no host synchronisation: capturable in a hipGraph, also on the first
 * execute.  A plan executes on one stream at a time. */
mxr_status mxr_ragged_execute(mxr_ragged_plan plan, hipStream_t stream);
/* The block size chosen for job `job`. */
mxr_status mxr_ragged_plan_block_size(mxr_ragged_plan plan, int job, int *M);
mxr_status mxr_ragged_plan_destroy(mxr_ragged_plan plan);

/* ------------------------------------------------------------------ CFAR */
/* Cell-averaging CFAR with a rectangular training window of win_y x win_x
 * cells minus a centred hole of guard_y x guard_x cells (CUT + guards), zero
 * padded at the edges, Pfa-derived threshold:
 *   sum = masked box sum; norm = in-bounds masked cell count;
 *   ba = sum/norm; alpha = norm*(pfa^(-1/norm) - 1); det = pow > alpha*ba.
 * MatX's mask (simple_radar_pipeline.h:237-241) is win 5x13, guard 3x3.
 * Replaces conv2d(xPow, mask, FULL) + trim + divide + calcDets
 * (simple_radar_pipeline.h:394-436, op body :86-98).
 * in: [C, Ny, Nx] f32 power, or c64 (|.|^2 taken on load, = abs2(tpcView)).
 * dets: [C, Ny, Nx] i32.  ba_out: [C, Ny, Nx] f32 or NULL (any strides, e.g.
 * MatX's trimmed slice of the padded ba tensor).
 * pow_out: [C, Ny, Nx] f32 or NULL (writes |in|^2 when in is c64). */
typedef struct mxr_cfar_params {
  int32_t win_y, win_x;     /* odd */
  int32_t guard_y, guard_x; /* odd, <= win */
  float pfa;
} mxr_cfar_params;
mxr_status mxr_cfar(const mxr_tensor *dets, const mxr_tensor *ba_out,
                    const mxr_tensor *pow_out, const mxr_tensor *in,
                    const mxr_cfar_params *params, hipStream_t stream);

/* Pulse-axis pass and CA-CFAR fused: canceller (tpc_mode as mxr_doppler),
 * window, nd-point Doppler FFT, |.|^2 and the CFAR of mxr_cfar in one
 * launch; the power map stays on chip unless pow_out is given.
 * y: [C, P, R] c64.  dets: [C, nd, R] i32; ba_out, pow_out: [C, nd, R] f32
 * or NULL.  Replaces ThreePulseCanceller + DopplerProcessing + CFARDetections
 * (examples/simple_radar_pipeline.h:321-357,394-436).  Fused for nd <= 128,
 * unit range strides and the window shapes instantiated in the library;
 * other cases run mxr_doppler then mxr_cfar with identical results (through
 * a stream-ordered temporary when pow_out is NULL). */
mxr_status mxr_doppler_cfar(const mxr_tensor *dets, const mxr_tensor *ba_out,
                            const mxr_tensor *pow_out, const mxr_tensor *y, int tpc_mode,
                            int64_t win_rows, int64_t nd, const mxr_cfar_params *params,
                            hipStream_t stream);

/* ----------------------------------------------------------- elementwise */
/* Fused elementwise expression: out[i] = f(in_0[i], ..., in_{n-1}[i], i) for
 * every index i of out, in one kernel -- the (lhs = expr).run(exec) path of
 * cudaExecutor::Exec over matxOpT*Kernel (include/matx/executors/cuda.h:
 * 84-230, include/matx/executors/kernel.h:41-222).  f is a postfix program
 * over complex float values (real inputs enter with a zero imaginary part;
 * a f32 output keeps the real part, an i32 output the truncated real part).
 * Every input has out's rank; broadcasting is a zero stride (MatX Clone,
 * include/matx/core/tensor.h:1101-1114).  Up to 8 inputs, 48 instructions,
 * 16 constants, stack depth 8. */
typedef enum mxr_ew_opcode {
  MXR_EW_LOAD = 1,     /* push in[a] at the current index                  */
  MXR_EW_CONST = 2,    /* push constant a (consts[2a] + i*consts[2a+1])     */
  MXR_EW_HAMMING = 3,  /* push 0.54-0.46cos(2 pi i_a/(b-1)) in float,
                          include/matx/generators/hamming.h:96-99           */
  MXR_EW_INDEX = 4,    /* push the index along dim a                        */
  MXR_EW_ADD = 10, MXR_EW_SUB = 11, MXR_EW_MUL = 12, MXR_EW_DIV = 13,
  MXR_EW_POW = 14,     /* real part: x^y (cuda::std::pow of calcDets)       */
  MXR_EW_GT = 15, MXR_EW_LT = 16, MXR_EW_GE = 17, MXR_EW_LE = 18,
  MXR_EW_EQ = 19,      /* real-part comparisons -> 1 or 0                   */
  MXR_EW_MAX = 20, MXR_EW_MIN = 21,
  MXR_EW_CONJ = 30, MXR_EW_ABS2 = 31, MXR_EW_ABS = 32, MXR_EW_SQRT = 33,
  MXR_EW_NEG = 34, MXR_EW_REAL = 35, MXR_EW_IMAG = 36, MXR_EW_EXP = 37,
  MXR_EW_LOG = 38, MXR_EW_COS = 39, MXR_EW_SIN = 40,
  MXR_EW_SELECT = 50   /* pop c, y, x: push (re c != 0) ? x : y             */
} mxr_ew_opcode;
/* POW, GT, LT, GE, LE, MAX, MIN, SQRT, COS, SIN and SELECT's condition take
 * real operands only: a program that feeds them a complex value (a complex
 * input or constant, or arithmetic on one) is rejected with
 * MXR_INVALID_TYPE before any launch. */
typedef struct mxr_ew_instr {
  int32_t op, a, b;
} mxr_ew_instr;
mxr_status mxr_ew_eval(const mxr_tensor *out, const mxr_tensor *const *ins, int nin,
                       const mxr_ew_instr *prog, int nprog, const float *consts,
                       int nconst, hipStream_t stream);

/* Sum of all elements of `in` (c64 or f32) into the one-element `out` (same
 * dtype), in float with a two-level tree (replaces sum_impl / cub_sum,
 * include/matx/transforms/reduce.h:646-655, transforms/cub.h:1828-1860). */
mxr_status mxr_sum(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream);

/* mxr_sum without any allocation.  No counterpart in MatX: it is for callers
 * that cannot allocate (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and result as
 * mxr_sum (bit-identical); the 512 per-workgroup partials (4 KB) live in a
 * file-scope device array that each GPU holds until the process ends (the
 * first launch caps its grid at 512 workgroups, so any input size fits).
 * Every call on one GPU shares that array: calls that overlap in time on
 * one GPU (e.g. on two streams) race and must be serialised by the caller. */
mxr_status mxr_sum_noalloc(const mxr_tensor *out, const mxr_tensor *in, hipStream_t stream);

/* --------------------------------------------------------- convolution */
/* Direct 1-D convolution along the last axis, batched over the leading
 * dims (filter rank 1 is broadcast, as conv1d_impl clones it,
 * include/matx/transforms/conv.h:467-532).  in/out c64 or f32, filter f32 or
 * c64; output length by mode (FULL n+L-1, SAME n, VALID n-L+1), SAME start
 * as include/matx/kernels/conv.cuh:155-168.  Any strides. */
mxr_status mxr_conv1d(const mxr_tensor *out, const mxr_tensor *in,
                      const mxr_tensor *filt, int mode, hipStream_t stream);

/* 1-D convolution (corr = 0) or correlation (corr = 1: the filter
 * reverse(conj(filt)), corr_impl include/matx/transforms/corr.h:60-67) by
 * the FFT method, matxFFTConv1DInternal include/matx/transforms/conv.h:51-172:
 * windows FULL [0, n+L-1), SAME [s, s+n) with s = L odd ? (L-1)/2 : L/2-1,
 * VALID [L-1, n).  in/out c64 or f32 (out c64 iff in or filt is), filt [L]
 * broadcast over the leading dims, any strides.  n + L - 1 <= 16384: one
 * fused kernel per row (power-of-two transform >= n + L - 1 in LDS).
 * Longer rows: c64, unit stride along the axis and L <= 8193 (overlap-save
 * correlation blocks for the outputs at FULL index >= L-1, the fused kernel
 * for the ones before). */
mxr_status mxr_conv1d_fft(const mxr_tensor *out, const mxr_tensor *in,
                          const mxr_tensor *filt, int mode, int corr, hipStream_t stream);

/* Direct 2-D convolution over the last two axes, batched over leading dims
 * (include/matx/transforms/conv.h:547-595; index rule :340-380). */
mxr_status mxr_conv2d(const mxr_tensor *out, const mxr_tensor *in,
                      const mxr_tensor *filt, int mode, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MXR_H_ */
