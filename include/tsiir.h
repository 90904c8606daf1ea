/*
 * tsiir.h - C ABI of the MI355X (gfx950) IIR filtering library.
 *
 * This is the drop-in boundary for torchscience's IIR filtering hot path.
 * The reference implements the path in Python over ATen ops:
 *   sosfilt              src/torchscience/signal_processing/filter/_sosfilt.py:13-133
 *   lfilter (DF2T core)  src/torchscience/signal_processing/filter/_lfilter.py:12-178
 *   batched_filter_apply src/torchscience/signal_processing/filter/_batched.py:449-566
 * and has no native entry point for it, so each function below replaces the
 * Python loop it names: a ctypes (or any C FFI) caller hands over device
 * pointers, sizes and a HIP stream, and the library launches hand-written HIP
 * kernels on that stream.
 *
 * Conventions
 *  - Plain C: no HIP, torch or C++ types.  `stream` is a hipStream_t passed
 *    as void* (NULL = the default stream).
 *  - All data pointers are device pointers.  The library never allocates,
 *    frees or synchronises; it launches asynchronously and keeps no pointer
 *    after returning.
 *  - Errors are returned as tsiir_status; nothing is thrown across the ABI.
 *  - Every call is reentrant.  Calls that share one workspace must not run
 *    concurrently (they would race on its scan status words).
 *
 * Data layout
 *  - Signals are (channels, samples) with unit stride along time and a row
 *    stride in elements between channels.
 *  - Channel c of the output reads input row (c % x_rows), filter
 *    (c / chans_per_filter) of `sos`, and initial-state row c of `zi`
 *    (row 0 when TSIIR_ZI_BROADCAST is set).
 *  - sos is (n_filters, n_sections, 6) contiguous, rows [b0 b1 b2 a0 a1 a2];
 *    each row is normalised by a0 when a0 != 1, as _lfilter.py:96-98 does.
 *  - zi / zf are (channels, n_sections, 2) contiguous, in the compute type
 *    (float for TSIIR_F32 / TSIIR_BF16 I/O, double for TSIIR_F64), or in
 *    double with TSIIR_STATE_F64.
 */
#ifndef TSIIR_H
#define TSIIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSIIR_ABI_VERSION 1

typedef enum tsiir_status {
  TSIIR_OK = 0,
  TSIIR_EINVAL = 1,        /* bad descriptor, pointer or size */
  TSIIR_EUNSUPPORTED = 2,  /* valid request this build does not implement */
  TSIIR_ELAUNCH = 3,       /* a kernel launch failed */
  TSIIR_EWORKSPACE = 4,    /* workspace missing or too small */
  TSIIR_ESCRATCH = 5,      /* needs more than the fixed scratch of
                              tsiir_sosfilt_forward_noalloc */
  TSIIR_EAGAIN = 6         /* serving queue full: poll, then submit again */
} tsiir_status;

typedef enum tsiir_dtype {
  TSIIR_F32 = 0,
  TSIIR_F64 = 1,
  TSIIR_BF16 = 2
} tsiir_dtype;

/* Arithmetic form of the recurrence (the compute type follows io_dtype). */
typedef enum tsiir_mode {
  TSIIR_MODE_PLAIN = 0,        /* DF2T in the compute type                   */
  TSIIR_MODE_COMPENSATED = 1   /* FP32/BF16 storage, FP64 recurrence updates */
} tsiir_mode;

/* flags */
#define TSIIR_HAS_ZI 0x1u        /* zi holds the initial state             */
#define TSIIR_WRITE_ZF 0x2u      /* write the final state to zf            */
#define TSIIR_ZI_BROADCAST 0x4u  /* one (n_sections, 2) zi for every channel */
#define TSIIR_REVERSE 0x8u       /* run along decreasing time (adjoint)    */
/* Forward work decomposition (default: chosen from the shape).  WINDOW:
 * windows of a channel run in parallel, chained by a look-back over scan
 * status words.  CHANNEL: one wave per channel walks its windows in order
 * (for at least as many channels as the GPU holds waves).  At most one. */
#define TSIIR_PATH_WINDOW 0x10u
#define TSIIR_PATH_CHANNEL 0x20u
/* FP32 recurrences (FP32 / BF16 I/O, plain mode) in the reference's DF2T
 * update order with FP64 carries between windows and waves, in place of the
 * default direct-form-I kernel whose FP32 lane scan loses accuracy on poles
 * very close to the unit circle.  Callers set it for such filters (the
 * Python layer does for any pole radius above 0.995). */
#define TSIIR_FP32_DF2T 0x80u
/* Testing only: every look-back wait of the forward gives up at once, as if
 * its spin limit had passed (see tsiir_workspace_set_error_mailbox): the
 * windows after the first of each channel come out NaN and the error is
 * reported, with every kernel exiting normally. */
#define TSIIR_DEBUG_GIVE_UP 0x100u
/* zi / zf hold FP64 values whatever io_dtype is (a caller that carries a
 * stream's state between chunks in FP64 with FP32 or BF16 samples). */
#define TSIIR_STATE_F64 0x200u
/* Few channels of long rows (fewer than 1024 channels, one input row per
 * channel, rows of at least 1024 / channels segments of 16384 samples, at
 * most 16 sections, no saved section outputs) take a segmented
 * form when every filter forgets its state within TSIIR_SEGMENT_WARMUP
 * samples: the infinity norm of Phi^4096 below 2^-64, Phi being the 2S x 2S
 * state-transition matrix of the cascade in DF2T form, checked on the device
 * in the call.  Each segment then runs as an independent row of the
 * many-channel kernels, starting from the state that a
 * TSIIR_SEGMENT_WARMUP-sample pass from zero state over the end of the
 * segment before it reaches, so no look-back wait links segments.  The
 * result equals the look-back form's to rounding.  The look-back form is
 * launched as well and runs only when the check fails.  This flag keeps the
 * look-back form only (for comparisons). */
#define TSIIR_NO_SEGMENT 0x400u
#define TSIIR_SEGMENT_WARMUP 4096

typedef struct tsiir_sosfilt_desc {
  int64_t channels;          /* output channels B                          */
  int64_t n_samples;         /* T                                          */
  int32_t n_sections;        /* S (>= 1)                                   */
  int32_t io_dtype;          /* tsiir_dtype of x and y                     */
  int32_t coef_dtype;        /* tsiir_dtype of sos (F32 or F64)            */
  int32_t mode;              /* tsiir_mode                                 */
  uint32_t flags;            /* TSIIR_* flags                              */
  int32_t reserved;
  int64_t x_rows;            /* distinct input rows (channel c reads c % x_rows) */
  int64_t x_row_stride;      /* elements between input rows               */
  int64_t y_row_stride;      /* elements between output rows              */
  int64_t chans_per_filter;  /* channels that share one sos filter (>= 1) */
} tsiir_sosfilt_desc;

/* Library ABI version (TSIIR_ABI_VERSION of the build). */
int tsiir_version(void);

/* Static text for a status code. */
const char* tsiir_status_string(tsiir_status status);

/* Device-memory workspace the forward call needs for this descriptor.
 * Replaces nothing in the reference (its Python loop allocates per step);
 * callers size one buffer with it and reuse it. */
size_t tsiir_sosfilt_workspace_bytes(const tsiir_sosfilt_desc* desc);

/* Prepare a workspace for first use: zero-fills it on `stream`.  Needed once
 * per allocation; later calls keep it consistent themselves. */
tsiir_status tsiir_workspace_init(void* workspace, size_t workspace_bytes,
                                  void* stream);

/* Error reporting of a workspace.  A look-back wait (a window waiting for
 * the state published by the window before it) that exceeds its spin limit
 * gives up, so that every kernel terminates: it sets a bit in the workspace
 * header's error word, writes the bit to the registered mailbox, and makes
 * the affected windows' outputs (and those that depend on them) NaN.  The
 * mailbox is a 32-bit word the kernels can write: pinned host memory
 * (hipHostMalloc; pass its host address) or device memory; NULL unregisters
 * it.  A caller polls the mailbox without synchronising.  The header is
 * zeroed by tsiir_workspace_init, so register after it (stream-ordered).
 * Bits 1 and 2 are look-back waits; tsiir_sosfiltfilt reports a filter
 * without a steady state the same way (TSIIR_ERR_STEADY_STATE).
 * No counterpart in the reference (its Python loop cannot time out). */
tsiir_status tsiir_workspace_set_error_mailbox(void* workspace, size_t workspace_bytes,
                                               uint32_t* mailbox, void* stream);

/* y = sosfilt(sos, x, zi) along time; zf = final state (when requested).
 * Replaces the section loop _sosfilt.py:117-128 and the per-sample DF2T loop
 * _lfilter.py:153-166 it calls, and the per-filter loop of
 * batched_filter_apply(filter_format="sos") _batched.py:513-531 (via
 * x_rows / chans_per_filter).  Also serves lfilter for filters of order <= 2
 * (a single section), _lfilter.py:12-178. */
tsiir_status tsiir_sosfilt_forward(const tsiir_sosfilt_desc* desc,
                                   const void* x, void* y, const void* sos,
                                   const void* zi, void* zf,
                                   void* workspace, size_t workspace_bytes,
                                   void* stream);

/* Zero-phase filtering: y = sosfiltfilt(sos, x) of the rows desc describes
 * (no zi / zf: desc->flags may hold only TSIIR_FP32_DF2T).  The rows are
 * extended at both ends by padlen samples (padtype 1 odd, 2 even, 3
 * constant; 0 none, with padlen 0), filtered forward from zi_unit * (first
 * extended sample), then backward from zi_unit * (last sample of the first
 * pass); y is the central n_samples of the second pass.  zi_unit: FP64
 * (n_filters, n_sections, 2), the steady state sosfilt_zi of each filter.
 * The extensions are never stored: two thread-per-channel kernels filter
 * them and hand FP64 states to the two sosfilt passes over the rows.
 * `scratch` (tsiir_sosfiltfilt_scratch_bytes, any device memory, no
 * initialisation) holds the first pass's rows; `workspace` is the sosfilt
 * workspace of desc (tsiir_sosfilt_workspace_bytes), used by both passes.
 * At most 8 sections and padlen * n_sections <= 16384, else
 * TSIIR_EUNSUPPORTED.  Replaces _sosfiltfilt.py:96-173 (the padded
 * concatenation, the two sosfilt calls and the flips).
 *
 * Without a copy of the coefficients on the host: padlen =
 * TSIIR_PADLEN_DEFAULT takes each filter's default of _sosfiltfilt.py:100-110
 * (3 * (2 n_sections + 1 - min(#b2 == 0, #a2 == 0))), which has to be below
 * n_samples whatever the coefficients (3 * (2 n_sections + 1) < n_samples,
 * else TSIIR_EINVAL); zi_unit = NULL forms sosfilt_zi on the device (equal
 * bit for bit to the FP64 closed form of _sosfilt_zi.py:55-86 with
 * _lfilter_zi.py:100-128).  A filter with no steady state there (a section
 * whose system is singular: 1 + a1 / a0 + a2 / a0 == 0, or sum(a) == 0 when
 * a0 == 0; the reference's solve raises) gives NaN outputs and sets
 * TSIIR_ERR_STEADY_STATE in the workspace's error word and mailbox
 * (tsiir_workspace_set_error_mailbox). */
#define TSIIR_PAD_NONE 0
#define TSIIR_PAD_ODD 1
#define TSIIR_PAD_EVEN 2
#define TSIIR_PAD_CONSTANT 3
#define TSIIR_PADLEN_DEFAULT (-1)
#define TSIIR_ERR_STEADY_STATE 4u
size_t tsiir_sosfiltfilt_scratch_bytes(const tsiir_sosfilt_desc* desc, int64_t padlen);
tsiir_status tsiir_sosfiltfilt(const tsiir_sosfilt_desc* desc, int padtype, int64_t padlen,
                               const void* x, void* y, const void* sos, const double* zi_unit,
                               void* scratch, size_t scratch_bytes,
                               void* workspace, size_t workspace_bytes, void* stream);

/* tsiir_sosfilt_forward without a workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * and harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The workspace is a fixed device array of the library,
 * TSIIR_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on a GPU
 * until the process ends; descriptors whose workspace
 * (tsiir_sosfilt_workspace_bytes) exceeds it return TSIIR_ESCRATCH.  The
 * call allocates and frees nothing.  The first call on a GPU zero-fills that
 * GPU's array on `stream` (make it outside graph capture).  Every call on a
 * GPU shares its array: calls that overlap in time on one GPU race, and the
 * caller must order them (one stream, or events).  Same kernels and results
 * as tsiir_sosfilt_forward. */
#define TSIIR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
tsiir_status tsiir_sosfilt_forward_noalloc(const tsiir_sosfilt_desc* desc,
                                           const void* x, void* y, const void* sos,
                                           const void* zi, void* zf, void* stream);

/* y = lfilter(b, a, x, zi) for a filter of order N = desc->n_sections
 * (1..64), coefficients shared by every channel.  `coef` holds b[0..N] then
 * a[0..N] (2N+2 values, zero-padded to a common length as
 * _lfilter.py:114-117 does) in coef_dtype == io_dtype (F32 or F64); they are
 * normalised by a[0] when a[0] != 1.  zi / zf are (channels, N).
 * Replaces the per-sample DF2T loop _lfilter.py:153-166 for orders above 2
 * (orders <= 2 go through tsiir_sosfilt_forward as one section).
 * TSIIR_REVERSE runs along decreasing time (zi the state at the end, zf
 * after the first sample), as for tsiir_sosfilt_forward. */
tsiir_status tsiir_lfilter_forward(const tsiir_sosfilt_desc* desc, const void* coef,
                                   const void* x, void* y, const void* zi, void* zf,
                                   void* stream);

/* tsiir_lfilter_forward with a workspace, which lets orders up to 16 run in
 * time-parallel form: chunks of 256 samples filtered from zero state, the
 * chunk start states scanned in FP64 across the chunks, then every chunk
 * filtered again from its start state.  Workspace: channels x chunks x order
 * FP64 values (tsiir_lfilter_workspace_bytes; 0 when the serial form is
 * used, for orders above 16 or fewer than 512 samples).  Same results to
 * rounding; same reference loop replaced. */
size_t tsiir_lfilter_workspace_bytes(const tsiir_sosfilt_desc* desc);
tsiir_status tsiir_lfilter_forward_ws(const tsiir_sosfilt_desc* desc, const void* coef,
                                      const void* x, void* y, const void* zi, void* zf,
                                      void* workspace, size_t workspace_bytes, void* stream);

/* Gradients of sosfilt (FP32 or FP64 compute).  Given dy (and dzf when zf
 * was an output), writes dx, dsos (n_filters, n_sections, 6) and dzi
 * (channels, n_sections, 2).  Replaces the autograd tape the reference
 * records through _lfilter.py:153-166 (gradients reach x, b0..b2, a1, a2 and
 * zi; the a0 column receives 0 when a0 == 1, as in the reference).
 * Requires x_rows == channels.  Coefficient sums are accumulated per window
 * and channel in the workspace and reduced per (filter, section).  dzf,
 * dsos and dzi may be NULL when not given / not wanted. */
size_t tsiir_sosfilt_backward_workspace_bytes(const tsiir_sosfilt_desc* desc);
tsiir_status tsiir_sosfilt_backward(const tsiir_sosfilt_desc* desc,
                                    const void* x, const void* sos,
                                    const void* zi, const void* dy,
                                    const void* dzf, void* dx, void* dsos,
                                    void* dzi, void* workspace,
                                    size_t workspace_bytes, void* stream);

/* Training form of tsiir_sosfilt_forward: also writes the output of every
 * section but the last to `saved`, (n_sections - 1, channels, n_samples)
 * contiguous in the io dtype (NULL allowed when n_sections == 1), so that
 * tsiir_sosfilt_backward_saved needs no forward recompute.  Replaces the
 * activations the reference's autograd tape keeps for _lfilter.py:153-166. */
tsiir_status tsiir_sosfilt_forward_saved(const tsiir_sosfilt_desc* desc,
                                         const void* x, void* y, const void* sos,
                                         const void* zi, void* zf, void* saved,
                                         void* workspace, size_t workspace_bytes,
                                         void* stream);

/* tsiir_sosfilt_backward from the forward output y (channels, n_samples) and
 * the section outputs `saved` written by tsiir_sosfilt_forward_saved, without
 * recomputing the forward.  Same gradients and conventions. */
size_t tsiir_sosfilt_backward_saved_workspace_bytes(const tsiir_sosfilt_desc* desc);
tsiir_status tsiir_sosfilt_backward_saved(const tsiir_sosfilt_desc* desc,
                                          const void* x, const void* y,
                                          const void* saved, const void* sos,
                                          const void* zi, const void* dy,
                                          const void* dzf, void* dx, void* dsos,
                                          void* dzi, void* workspace,
                                          size_t workspace_bytes, void* stream);

/* ---- FIR filtering ----------------------------------------------------------
 * y[n] = sum_{j<K} h[j] x[n + shift - j] for n in [0, n_samples), with x
 * taken as zero outside [0, x_len).  Channel c reads input row c % x_rows and
 * tap row c / chans_per_filter of `taps` (n_filters, n_taps) contiguous; all
 * data (taps, x, y, zi, zf) is in io_dtype (F32 or F64). */
#define TSIIR_FIR_REVERSE_TAPS 0x40u  /* use h[K-1-j] in place of h[j]       */

typedef struct tsiir_fir_desc {
  int64_t channels;          /* output channels B                          */
  int64_t n_samples;         /* output samples per channel T               */
  int64_t x_len;             /* input samples per row                      */
  int32_t n_taps;            /* K (1 .. 2^20)                              */
  int32_t io_dtype;          /* tsiir_dtype: F32 or F64                    */
  int32_t shift;             /* output n reads inputs n + shift - j        */
  uint32_t flags;            /* HAS_ZI, WRITE_ZF, ZI_BROADCAST, FIR_REVERSE_TAPS */
  int64_t x_rows;            /* distinct input rows                        */
  int64_t x_row_stride;      /* elements between input rows               */
  int64_t y_row_stride;      /* elements between output rows              */
  int64_t chans_per_filter;  /* channels that share one tap row (>= 1)     */
} tsiir_fir_desc;

/* FIR filtering.  With shift = 0, x_len = n_samples and the lfilter state
 * flags it is lfilter(b, [a0], x, zi) for taps b / a0: zi / zf are
 * (channels, K - 1), zi is added to the first K - 1 outputs and zf is the
 * final DF2T state (the outputs K - 1 samples past the end).  Replaces the
 * per-sample DF2T loop _lfilter.py:153-166 for a = [a0] (with b normalised as
 * _lfilter.py:96-98 does), the per-filter FIR loop of batched_filter_apply
 * _batched.py:533-561 (via x_rows / chans_per_filter), and the FFT
 * convolution of fftfilt _fftfilt.py:60-92 (shift = (K - 1) / 2, no state).
 * The state flags require shift == 0. */
tsiir_status tsiir_fir_forward(const tsiir_fir_desc* desc, const void* taps, const void* x,
                               void* y, const void* zi, void* zf, void* stream);

/* FFT (overlap-save) form of tsiir_fir_forward: taps (n_filters, K), one
 * filter per chans_per_filter consecutive channels (channels must be a
 * multiple of chans_per_filter unless chans_per_filter >= channels, one
 * filter), no state flags (HAS_ZI / WRITE_ZF and a ragged last filter give
 * TSIIR_EINVAL), 1 <= K <= 2048 (else TSIIR_EUNSUPPORTED); honours shift,
 * x_rows, the strides and TSIIR_FIR_REVERSE_TAPS.  Same outputs up to
 * rounding; its cost per output does not grow with K.  Replaces the
 * rfft / multiply / irfft of fftfilt _fftfilt.py:76-91.  The workspace (size
 * from tsiir_fir_fft_workspace_bytes, 0 for an invalid descriptor) holds the
 * twiddles and one transform per filter (32 KB FP32, 64 KB FP64 each); it is rewritten on every call, so calls
 * that share one workspace must be ordered on one stream. */
size_t tsiir_fir_fft_workspace_bytes(const tsiir_fir_desc* desc);
tsiir_status tsiir_fir_forward_fft(const tsiir_fir_desc* desc, const void* taps, const void* x,
                                   void* y, void* workspace, size_t workspace_bytes,
                                   void* stream);

/* Gradient of the taps: dtaps[f][j] = sum over the channels c of filter f of
 * sum_{n < n_samples} g[c][n] x[c % x_rows][n + shift - j], x zero outside
 * [0, x_len); g has row stride y_row_stride.  The input gradient needs no
 * entry of its own: it is tsiir_fir_forward of g with TSIIR_FIR_REVERSE_TAPS
 * and shift K - 1 - shift.  Together they replace the autograd tape the
 * reference records through _lfilter.py:153-166 (a = [a0]) and through
 * fftfilt's FFTs (_fftfilt.py:81-85).  Requires channels % chans_per_filter
 * == 0 and fewer than 65536 filters; sums are accumulated in FP64. */
size_t tsiir_fir_backward_taps_workspace_bytes(const tsiir_fir_desc* desc);
tsiir_status tsiir_fir_backward_taps(const tsiir_fir_desc* desc, const void* x, const void* g,
                                     void* dtaps, void* workspace, size_t workspace_bytes,
                                     void* stream);

/* ---- analysis ---------------------------------------------------------------
 * Frequency response of n_filters cascades of n_sections second-order
 * sections: out[f][k] = prod_s (b0 + b1 z^-1 + b2 z^-2) / (a0 + a1 z^-1 +
 * a2 z^-2) at z = exp(j w[k]), as interleaved FP64 (re, im) pairs,
 * (n_filters, n_freq, 2).  sos is (n_filters, n_sections, 6) contiguous in
 * coef_dtype (F32 or F64); w holds n_freq angular frequencies (radians per
 * sample, FP64).  a0 is used as given.  Replaces the section loop of
 * frequency_response_sos _frequency_response_sos.py:128-137. */
tsiir_status tsiir_sos_frequency_response(const void* sos, int32_t coef_dtype,
                                          int64_t n_filters, int32_t n_sections,
                                          const double* w, int64_t n_freq, double* out,
                                          void* stream);

/* B(z^-1) / A(z^-1) at z = exp(j w[k]) for n_filters direct-form filters:
 * b (n_filters, nb), a (n_filters, na) FP64 (an empty polynomial is 1:
 * na = 0 makes an FIR),
 * out (n_filters, n_freq, 2) interleaved FP64 (re, im).  Both polynomials by
 * Horner's rule from the highest power.  Replaces the polynomial
 * evaluations of frequency_response _frequency_response.py:104-116 and
 * frequency_response_fir _frequency_response_fir.py:66-79. */
tsiir_status tsiir_ba_frequency_response(const double* b, int32_t nb, const double* a,
                                         int32_t na, int64_t n_filters, const double* w,
                                         int64_t n_freq, double* out, void* stream);

/* ---- batched IIR design -------------------------------------------------------
 * FP64 design of n_filters digital filters of one family, band type and
 * order into second-order sections, one cutoff (lowpass / highpass) or one
 * (low, high) pair (bandpass / bandstop) per filter, as fractions of the
 * Nyquist frequency.  Follows the reference's design chain (prototype,
 * pre-warp 4 tan(pi f / 2), band transform, bilinear map at fs = 2, its
 * pole/zero pairing and even gain split).  Replaces the per-filter loops of
 * batched_butterworth_design / batched_chebyshev_type_1_design /
 * batched_chebyshev_type_2_design _batched.py:22-344. */
#define TSIIR_DESIGN_BUTTERWORTH 0
#define TSIIR_DESIGN_CHEBYSHEV1 1   /* param: passband ripple in dB (> 0)     */
#define TSIIR_DESIGN_CHEBYSHEV2 2   /* param: stopband attenuation in dB (> 0) */
#define TSIIR_BAND_LOWPASS 0
#define TSIIR_BAND_HIGHPASS 1
#define TSIIR_BAND_BANDPASS 2
#define TSIIR_BAND_BANDSTOP 3

typedef struct tsiir_iir_design_desc {
  int64_t n_filters;
  int32_t family;       /* TSIIR_DESIGN_*                               */
  int32_t band;         /* TSIIR_BAND_*                                 */
  int32_t order;        /* 1 .. 32                                      */
  int32_t out_dtype;    /* F32 or F64                                   */
  double param;         /* ripple / attenuation in dB (Chebyshev)       */
} tsiir_iir_design_desc;

/* Sections per filter: ceil(order / 2), or order for band filters (0 for
 * an unsupported order). */
int32_t tsiir_iir_design_sections(const tsiir_iir_design_desc* desc);

/* cutoffs: (n_filters) or (n_filters, 2) FP64 device array, each in (0, 1)
 * (checked by the caller); sos: (n_filters, sections, 6) device output in
 * out_dtype, rows [b0 b1 b2 1 a1 a2]. */
tsiir_status tsiir_iir_design_sos(const tsiir_iir_design_desc* desc, const double* cutoffs,
                                  void* sos, void* stream);

/* tsiir_iir_design_sos that also writes the Jacobian of every coefficient
 * with respect to the filter's cutoffs: jacobian (n_filters, sections, 6, 2)
 * FP64, the last index the low (or only) and the high cutoff (0 for a
 * single cutoff).  Replaces the autograd the reference records through its
 * design ops (the cutoff gradients of _batched.py:22-344). */
tsiir_status tsiir_iir_design_sos_jacobian(const tsiir_iir_design_desc* desc,
                                           const double* cutoffs, void* sos,
                                           double* jacobian, void* stream);

/* ---- serving: ragged batches of streaming chunks --------------------------
 * The reference has no serving layer; these entry points serve the streaming
 * use of sosfilt with zi/zf (_sosfilt.py:98-133): a chunk of a persistent
 * stream is filtered from the stream's carried state, and zf becomes the
 * state for the next chunk.  One launch filters a whole batch of chunks of
 * different streams, channel counts, lengths and section counts. */

/* One channel of one chunk.  x / y: n_samples FP32 samples (device).  table:
 * the stream filter's coefficient table (tsiir_serve_prepare_filter).
 * state: (n_sections, 2) FP64 state of this channel, read as zi and
 * overwritten with zf.  n_sections in 1..32. */
typedef struct tsiir_serve_item {
  const void* x;
  void* y;
  const double* table;
  double* state;
  int64_t n_samples;
  int32_t n_sections;
  int32_t reserved;
} tsiir_serve_item;

/* Bytes of one filter's coefficient table (0 if n_sections is out of range). */
size_t tsiir_serve_table_bytes(int32_t n_sections);

/* Build a filter's table on the device from its (n_sections, 6) sos
 * (coef_dtype F32 or F64); rows are normalised by a0 when a0 != 1. */
tsiir_status tsiir_serve_prepare_filter(const void* sos, int32_t coef_dtype, int32_t n_sections,
                                        double* table, void* stream);

/* Filter every item of a batch (items: device array of n_items).  Items must
 * not share a state slot (one chunk per stream and channel per batch); an
 * item's result does not depend on the other items.  mode: TSIIR_MODE_PLAIN
 * (FP32 recurrence) or TSIIR_MODE_COMPENSATED (FP32 storage, FP64 updates). */
tsiir_status tsiir_serve_batch(const tsiir_serve_item* items, int32_t n_items, int32_t mode,
                               void* stream);

/* ---- persistent serving launch -----------------------------------------
 * A queue of items (tsiir_serve_item) served by a persistent launch: a fixed
 * grid of waves that keeps taking submitted items while it runs, so a
 * serving loop submits chunks as they arrive instead of launching a kernel
 * per batch.  The queue owns a HIP stream, a ring of `capacity` item slots
 * and completion words in host memory the device reads and writes
 * coherently, and a device claim counter: tsiir_serve_create allocates them,
 * tsiir_serve_destroy frees them.  Each item is walked by one wave from its
 * FP64 state, which the item overwrites with the final state (as
 * tsiir_serve_batch); an item's result does not depend on the other items.
 * The launch drains by itself after idle_us microseconds without work and is
 * launched again by the next submit or poll that finds work waiting.
 * Submitted items must have their input ready (the caller synchronises the
 * work that produced it) and must not share a state slot with an item still
 * in flight.  One queue per thread, or the caller serialises the calls.
 * The reference has no serving layer; this serves the streaming use of
 * sosfilt with zi / zf (_sosfilt.py:98-133). */
typedef struct tsiir_serve_queue tsiir_serve_queue;

/* capacity: item slots (tickets in flight at once); workgroups: waves of the
 * persistent launch; mode: TSIIR_MODE_PLAIN or TSIIR_MODE_COMPENSATED. */
tsiir_status tsiir_serve_create(int32_t capacity, int32_t workgroups, int32_t mode,
                                int32_t idle_us, tsiir_serve_queue** out);

/* Append n_items items (host array, copied); *first_ticket gets the ticket of
 * the first (tickets are consecutive).  TSIIR_EAGAIN when the slots they
 * need still hold unfinished items (nothing is submitted then). */
tsiir_status tsiir_serve_submit(tsiir_serve_queue* queue, const tsiir_serve_item* items,
                                int32_t n_items, uint64_t* first_ticket);

/* *all_done = 1 when every ticket of [first_ticket, first_ticket + n_items)
 * has completed (outputs and states written and visible), else 0.  Never
 * blocks. */
tsiir_status tsiir_serve_poll(tsiir_serve_queue* queue, uint64_t first_ticket, int32_t n_items,
                              int32_t* all_done);

/* Lets the launch finish the submitted items, waits for it, frees the queue. */
tsiir_status tsiir_serve_destroy(tsiir_serve_queue* queue);

#ifdef __cplusplus
}
#endif

#endif /* TSIIR_H */
