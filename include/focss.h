/*
 * focss C-ABI -- the drop-in boundary between the focss C++ host API
 * (Field / Fiber / Amplifier / SSFM / link drivers) and the MI355X (gfx950)
 * kernels.  Plain C types only: pointers, sizes, PODs.  Every entry point
 * returns 0 on success or a negative FOCSS_E_* code; focss_last_error()
 * returns the message of the calling thread's last failure.
 *
 * Field layout everywhere: mode-major interleaved complex, element
 * (mode, sample) at data[sample + mode * samples]   (src/focss/field.cpp:140-155).
 * precision FOCSS_F64 = complex<double> (16 B), FOCSS_F32 = complex<float> (8 B).
 * samples must be a power of two, 2^6 .. 2^20 (2 .. 2^20 for focss_fft).
 *
 * A context owns one HIP stream on one device plus cached device buffers and
 * captured graphs.  A context is used by one host thread at a time; distinct
 * contexts may run concurrently.
 */
#ifndef FOCSS_H_
#define FOCSS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOCSS_OK 0
#define FOCSS_E_ARG -1      /* invalid argument (sizes, grid, null pointer) */
#define FOCSS_E_HIP -2      /* HIP runtime error */
#define FOCSS_E_UNSUPPORTED -3
#define FOCSS_E_NUMERIC -4  /* non-finite result detected */

#define FOCSS_FAULT_ALLOC 1 /* focss_inject_fault: device allocations */

#define FOCSS_F64 0
#define FOCSS_F32 1

#define FOCSS_UNIFORM 0     /* SSFM::UNIFORM     src/focss/solver/ssfm.cpp:9 */
#define FOCSS_LOGARITHMIC 1 /* SSFM::LOGARITHMIC src/focss/solver/ssfm.cpp:10 */
#define FOCSS_ADAPTIVE 2    /* SSFM::ADAPTIVE    src/focss/solver/ssfm.cpp:11 */

typedef struct focss_ctx focss_ctx;

/* focss::Fiber POD, SI units (src/focss/module/fiber.h:5-21) */
typedef struct {
    double alpha;      /* power attenuation [1/m] */
    double beta2;      /* [s^2/m] */
    double gamma;      /* [1/(W m)] */
    double length;     /* [m] */
    double wavelength; /* [m] */
    double kappa;      /* polarisation coupling */
} focss_fiber;

/* Grid parameters of focss::SSFM (src/focss/solver/ssfm.h:25-29) */
typedef struct {
    int grid;                /* FOCSS_UNIFORM / FOCSS_LOGARITHMIC / FOCSS_ADAPTIVE */
    int total_steps;         /* UNIFORM steps per span (set_total_steps) */
    double max_phase_shift;  /* set_maximum_phase_shift */
    double max_step_size;    /* set_maximum_step_size [m] */
} focss_grid;

/* One waveform's link program:
 *   for span in 0..n_spans-1:
 *       field *= span_pre_scale                (Amplifier::drop_power, amplifier.cpp:37-42)
 *       SSFM(fiber, grid).run(field, direction) (ssfm.cpp:39-60)
 *       field *= span_post_scale               (Amplifier::amplify gain, amplifier.cpp:15-18)
 *       if noise_sigma: field += noise_sigma * (z0 + i z1), z ~ N(0,1) from
 *                       Philox4x32-10(seed, stream = span, index = flat element)
 * This is forward_propagation / backward_propagation / cd_compensation
 * (src/tasks/systems.cpp:26-91) and SSFM::run itself (n_spans = 1). */
typedef struct {
    focss_fiber fiber;
    focss_grid grid;
    double direction;        /* +1 FORWARD, -1 BACKWARD (ssfm.cpp:13-14) */
    int n_spans;
    int modes;               /* 1 or 2 (Manakov) */
    int samples;             /* any length in [2, 2^19], or a power of two up to 2^20 */
    int pad;
    double dt;               /* sample spacing [s] (Domain::time_step) */
    double center_frequency; /* [Hz] (Domain::center_frequency) */
    double span_pre_scale;
    double span_post_scale;
    double noise_sigma;      /* per-quadrature sigma, 0 = no ASE */
    uint64_t noise_seed;
} focss_link;

/* Fault injection for tests of the error paths (no counterpart in the
 * reference, which has none): with kind FOCSS_FAULT_ALLOC the next `count`
 * device allocations the library makes fail as an exhausted device would
 * (FOCSS_E_HIP); count 0 clears it.  The failed call leaves its context
 * usable. */
int focss_inject_fault(int kind, int count);

/* ---- context / memory ---- */
int focss_ctx_create(int device, focss_ctx** out);
int focss_ctx_destroy(focss_ctx* ctx);
const char* focss_last_error(void);
int focss_device_alloc(focss_ctx* ctx, size_t bytes, void** dptr);
int focss_device_free(focss_ctx* ctx, void* dptr);
int focss_memcpy_h2d(focss_ctx* ctx, void* dst, const void* src, size_t bytes);
int focss_memcpy_d2h(focss_ctx* ctx, void* dst, const void* src, size_t bytes);
int focss_synchronize(focss_ctx* ctx);

/* ---- the hot path ----
 * A program whose field power stops being finite (a NaN or infinite sample
 * in the input, or an overflow on the way) stops at that point; the call
 * then returns FOCSS_E_NUMERIC naming the waveforms, after the others have
 * finished.
 * Runs the link programs of n_waves waveforms resident in device memory.
 * Waveform i occupies modes*samples complex elements at
 * (complex*)batch + offsets[i] (element offsets).  Waveforms may differ in
 * samples, modes, fibre, grid and span count; all run in one stream, the
 * small ones LDS-resident in one launch, the large ones as a hipGraph of
 * alternating time/spectral passes replayed until every waveform is done
 * (device-side step control, no host sync per step).
 * total_steps (optional, n_waves ints) receives the split steps performed.
 * Replaces SSFM::run (src/focss/solver/ssfm.cpp:39-60) and the span loops of
 * forward_propagation / backward_propagation / cd_compensation
 * (src/tasks/systems.cpp:26-91, src/tasks/wdm_run.cpp:26-47). */
int focss_link_run(focss_ctx* ctx, int precision, int n_waves, const focss_link* links,
                   void* batch, const int64_t* offsets, int* total_steps);

/* focss_link_run with the ADAPTIVE grid's step schedule recorded or
 * replayed, for precision studies (no counterpart in the reference):
 * schedule is a device array of n_waves x capacity doubles, waveform i's at
 * schedule + i * capacity.  FOCSS_SCHEDULE_RECORD writes each ADAPTIVE step
 * taken (its length, negated for a span's last step; steps beyond capacity
 * are not recorded); FOCSS_SCHEDULE_REPLAY takes those steps instead of
 * estimating them from the field's peak power (ssfm.cpp:92-118), so an FP32
 * run can follow the FP64 run's schedule exactly, and continues with its
 * own estimates past capacity.  Other grids are unaffected. */
#define FOCSS_SCHEDULE_RECORD 1
#define FOCSS_SCHEDULE_REPLAY 2
int focss_link_run_schedule(focss_ctx* ctx, int precision, int n_waves, const focss_link* links, void* batch,
                            const int64_t* offsets, int* total_steps, int mode, double* schedule, int capacity);

/* Measurement: runs only the first n_pairs split steps of the programs, one
 * kernel launch at a time, each bracketed by HIP events on the context's
 * stream.  avg_ms[0] = mean time-pass duration, avg_ms[1] = mean
 * spectral-pass duration (two-pass waveforms), avg_ms[2] = duration of the
 * LDS-resident launch (whole programs).  The field is left mid-program. */
int focss_link_profile(focss_ctx* ctx, int precision, int n_waves, const focss_link* links,
                       void* batch, const int64_t* offsets, int n_pairs, double* avg_ms);

/* Measurement: the link programs run as focss_link_run runs them (entry
 * pass, then replays of the captured graph of pass pairs), but stop after
 * n_replays replays; *ms_per_pair = the time between one event before the
 * first replay and one after the last, divided by the number of (spectral,
 * time) pass pairs replayed.  The batch is left part-way through its
 * programs. */
int focss_link_profile_graph(focss_ctx* ctx, int precision, int n_waves, const focss_link* links,
                             void* batch, const int64_t* offsets, int n_replays, double* ms_per_pair);

/* Host-only argument check of one link program (no device needed): the
 * guards the reference leaves to assert or to an endless loop -- h_max = 0
 * on the ADAPTIVE grid (ssfm.cpp:99), alpha = 0 on the LOGARITHMIC grid
 * (ssfm.cpp:76), non-power-of-two or out-of-range sizes, bad direction. */
int focss_link_validate(const focss_link* link);

/* Device-to-device copy by a plain streaming kernel (HBM counter calibration). */
int focss_device_copy(focss_ctx* ctx, void* dst, const void* src, size_t bytes);

/* The context's hipStream_t, for callers that record their own events. */
void* focss_ctx_stream(focss_ctx* ctx);

/* Host-buffer convenience for one waveform: complex<double> in/out, computed
 * at `precision` on the device.  Used by focss::SSFM::run. */
int focss_link_run_host(focss_ctx* ctx, int precision, const focss_link* link,
                        double* field_c128, int* total_steps);

/* Batched in-place DFT of `modes` rows of `samples` (device memory).
 * sign -1: forward e^{-2 pi i jk/N} then x 1/N (Field::fft_inplace,
 * src/focss/field.cpp:343-347); sign +1: inverse, unscaled
 * (Field::ifft_inplace, src/focss/field.cpp:349-352). */
int focss_fft(focss_ctx* ctx, int precision, int sign, int modes, int samples, void* data);

/* focss_fft without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as focss_fft; the intermediate of the two-pass lengths
 * (above 2^13) lives in a fixed 32 MiB buffer of the library, one per device,
 * held until the process ends, so modes x samples x element size must not
 * exceed 32 MiB (FOCSS_E_ARG otherwise).  Calls that overlap on one device
 * (several contexts or host threads) race on that buffer. */
int focss_fft_noalloc(focss_ctx* ctx, int precision, int sign, int modes, int samples, void* data);

/* Amplifier::add_noise (src/focss/module/amplifier.cpp:20-28) on a device
 * field: element i (flat, mode-major) += sigma * (z0 + i z1), (z0, z1) the
 * Box-Muller pair of Philox4x32-10 block (i, stream) under key seed. */
int focss_add_noise(focss_ctx* ctx, int precision, int modes, int samples, double sigma,
                    uint64_t seed, uint32_t stream, void* data);

/* Field::operator*=(real / complex) (src/focss/field.cpp:381-392) on a
 * device field: every element of the modes x samples field times
 * (re + i im).  Stream-ordered. */
int focss_field_scale(focss_ctx* ctx, int precision, int modes, int samples, double re, double im, void* data);

/* ---- perturbation-triplet equalizer ----
 * Weights W are (2R+1)^2 complex<double> (interleaved re, im), row-major:
 * W(m, n) at index (R + m) * (2R + 1) + (R + n).  Fields are device
 * buffers in the mode-major layout. */

/* SINO::cubic (src/focss/equalizer/sino.cpp:203-220): out(mode, k) =
 * sum_{|m|+|n|<=R} W(m,n) A_mode(k+m) sum_p A_p(k+n) conj(A_p(k+m+n)) for
 * R <= k < samples - R, 0 elsewhere. */
int focss_sino_cubic(focss_ctx* ctx, int precision, int radius, const double* weights, int modes,
                     int samples, const void* in, void* out);

/* SINO::train (src/focss/equalizer/sino.cpp:42-159), type 0 ORDINARY,
 * 1 FIXED (uses linear_weight_in), 2 HOMOGENEOUS.  tx, rx: complex<double>
 * device fields.  Folded symmetric features and the Gram matrix F^H F are
 * built on the device; the z-scored normal equations are solved by
 * Cholesky on the host.  Outputs (host): weights (2R+1)^2, intercept (2
 * doubles), linear weight (2 doubles). */
int focss_sino_train(focss_ctx* ctx, int radius, int type, const double* linear_weight_in, int modes,
                     int samples, const void* tx, const void* rx, double* weights, double* intercept,
                     double* linear_weight);

/* WdmPPE::spm and WdmPPE::xpm (src/focss/equalizer/wdm_sino.cpp:256-295) over
 * six envelopes a6 = [6][samples] (row = channel * 2 + polarisation, rows
 * 2..3 the channel of interest).  C is (2c+1)^2, D is (2d+1)^2; a negative
 * radius skips that term.  spm, xpm: [2][samples] device outputs. */
int focss_wdm_ppe(focss_ctx* ctx, int precision, int c_radius, const double* C, int d_radius,
                  const double* D, int samples, const void* a6, void* spm, void* xpm);

/* WdmPPE::train (src/focss/equalizer/wdm_sino.cpp:29-141): tx6, rx6 are
 * complex<double> device envelopes [6][samples] (the rows of load_field's
 * matrix for three dual-polarisation channels).  Rows pol-major over samples
 * [r, samples - r), r = max(c, d); features: the folded SPM set of radius c,
 * T1(m, n) + T3(m, -n) over the d-diamond, the COI sample.  Feature matrix
 * and F^H F on the device, z-scored normal equations by Cholesky on the
 * host.  Outputs (host, interleaved doubles): C (2c+1)^2 and D (2d+1)^2
 * (skipped for radius -1), intercept, linear weight. */
int focss_wdm_ppe_train(focss_ctx* ctx, int c_radius, int d_radius, int samples, const void* tx6,
                        const void* rx6, double* C, double* D, double* intercept, double* linear_weight);

/* ---- transmitter / receiver DSP on the device ----
 * Fields are mode-major device buffers [modes][samples] of the given
 * precision, samples a power of two <= 2^20.  Each call is stream-ordered
 * (focss_filter, focss_rrc_shaping and focss_rrc_sampling also synchronise
 * once to upload the impulse response).  The DFTs follow Field::fft_inplace
 * (forward x 1/N). */
/* frequency_shift: f *= exp(i 2 pi frequency dt s)   (modulation.cpp:255-264) */
int focss_frequency_shift(focss_ctx* ctx, int precision, int modes, int samples, double dt, double frequency,
                          void* data);
/* bandpass_filter: zero the bins cut .. N - cut, cut = frequency N dt + 1
 *                                                    (modulation.cpp:266-277) */
int focss_bandpass_filter(focss_ctx* ctx, int precision, int modes, int samples, double dt, double frequency,
                          void* data);
/* wdm_select: shift channel channel_index to baseband and bandpass it
 *                                                    (modulation.cpp:248-253) */
int focss_wdm_select(focss_ctx* ctx, int precision, int modes, int samples, double dt, int channel_index,
                     double channel_bandwidth, const void* in, void* out);
/* Field::decimate: spectral truncation to samples / factor (field.cpp:299-320);
 * out holds modes x samples / factor. */
int focss_decimate(focss_ctx* ctx, int precision, int modes, int samples, int factor, const void* in, void* out);
/* Field::filter_inplace with a host impulse response of `taps` complex128
 * values, centred at sample 0                        (field.cpp:358-372) */
int focss_filter(focss_ctx* ctx, int precision, int modes, int samples, int taps, const double* impulse_response,
                 void* data);
/* rrc_shaping: upsample symbols [modes][n_symbols] by osf, RRC filter,
 * x sqrt(gain) osf                                   (modulation.cpp:86-90) */
int focss_rrc_shaping(focss_ctx* ctx, int precision, int modes, int n_symbols, const void* symbols, int osf,
                      double roll_off, double gain, void* out);
/* rrc_sampling: RRC filter, take every osf-th sample of the flat buffer,
 * x sqrt(gain); out holds modes x samples / osf       (modulation.cpp:92-96) */
int focss_rrc_sampling(focss_ctx* ctx, int precision, int modes, int samples, const void* signal, int osf,
                       double roll_off, double gain, void* out);

/* ---- on-disk field records ----
 * save_field / load_field files (src/tasks/wdm_run.cpp:55-77,
 * src/tasks/wdm_scoring.cpp:79-103) hold four int32 {n_channels, n_samples,
 * n_modes, flags} and then complex<double> values in sample-major order
 * [sample][channel][mode].  These calls move values between that order (a
 * device copy of the file body, `records`) and the field layout,
 * envelope-major [channel * n_modes + mode][samples] in the given precision
 * (the rows of load_field's matrix).  envelopes in [1, 2^20]; in and out must
 * not alias.  Stream-ordered. */
int focss_field_from_records(focss_ctx* ctx, int precision, int envelopes, int samples, const void* records,
                             void* field);
int focss_field_to_records(focss_ctx* ctx, int precision, int envelopes, int samples, const void* field,
                           void* records);

/* ---- constant equalisers (src/focss/equalizer/linear.h) ----
 * desired and actual are device fields [modes][samples] in the given
 * precision; one device pass forms sum |d|^2, sum |a|^2 and sum conj(a) d in
 * FP64; each call synchronises and returns its estimate on the host. */
/* PhaseShiftEqualizer::estimate (phase_shift_equalizer.cpp:30-46): the angle
 * 2 pi i / angle_steps whose rotation of actual maximises q2_factor against
 * desired, the first maximum, 0 when no rotation beats the unrotated field. */
int focss_pse_estimate(focss_ctx* ctx, int precision, int modes, int samples, const void* desired, const void* actual,
                       int angle_steps, double* angle);
/* ScalarEqualizer::estimate (scalar_equalizer.cpp:22-33): sum conj(a) d / sum |a|^2,
 * written as scalar[0] + i scalar[1]. */
int focss_scalar_estimate(focss_ctx* ctx, int precision, int modes, int samples, const void* desired,
                          const void* actual, double* scalar);
/* HomothetyEqualizer::estimate (homothety_equalizer.cpp:22-33): |sum conj(a) d / sum |a|^2|. */
int focss_homothety_estimate(focss_ctx* ctx, int precision, int modes, int samples, const void* desired,
                             const void* actual, double* scalar);

/* ---- learned few-step DBP (no counterpart in the reference) ----
 * Step j: y = complex FIR (taps, circular, centre taps/2) of x, then
 * z = y exp(i phi), phi(t) = sum_m g_{j,m} |y(t - m + kerr_taps/2)|^2 (a real
 * filter on the intensity; kerr_taps = 1 is the plain Kerr step
 * exp(i theta_j |y|^2)).  params (device, float): the FIR taps [steps][taps]
 * as interleaved complex64, followed by g [steps][kerr_taps].  Fields are
 * complex64 device buffers [batch][modes][samples]; modes 2 is the Manakov
 * (dual-polarisation) model, one FIR per step applied to each polarisation
 * and the phase driven by the intensity summed over both.  samples is a
 * power of two in [1024, 16384 / modes], taps 5, 15, 23, 31, 47 or 63,
 * kerr_taps 1, 5, 9 or 15.  Loss =
 * (1/(modes samples batch)) sum_{b, mode, t} |z - target|^2.  Measured against
 * plain DBP (focss_link_run); callers are learned-equaliser training and
 * inference. */
size_t focss_ldbp_workspace_bytes(int steps, int modes, int samples, int batch);
/* Inference (out) and/or loss against target (*loss, host double; the call
 * then synchronises).  Without target the call is stream-ordered and
 * allocates nothing. */
int focss_ldbp_forward(focss_ctx* ctx, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                       const float* params, const void* in, void* out, const void* target, double* loss);
/* Inference on the matrix cores: the same model with the FIR of every step
 * as FP16 MFMA products accumulated in FP32 and the field held in FP16
 * between steps (about 3e-3 relative deviation from the FP32 path after 80
 * steps).  Stream-ordered; the context keeps a buffer for the per-step
 * matrix operands (grown, with a synchronisation, when a call needs more). */
int focss_ldbp_infer(focss_ctx* ctx, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                     const float* params, const void* in, void* out);
/* Accumulates into grad (device, double [2 steps taps + steps kerr_taps],
 * zeroed by the caller) the Wirtinger gradient dL/d(conj h) (re, im) and
 * dL/dg, and the loss into *loss (device double).  The backward pass runs one
 * launch per step over 512-sample chunks of every waveform and adds the
 * chunks' sums in a fixed order, so the gradient does not depend on
 * scheduling.  Stream-ordered, no host sync. */
int focss_ldbp_gradients(focss_ctx* ctx, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                         const float* params, const void* in, const void* target, void* workspace,
                         size_t workspace_bytes, double* grad, double* loss);
/* One training step: gradients then Adam (t = step number from 1, optional
 * per-parameter learning-rate scale).  *loss_host receives the batch loss
 * before the update. */
int focss_ldbp_train_step(focss_ctx* ctx, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                          float* params, float* adam_m, float* adam_v, int t, float lr, const float* lr_scale,
                          const void* in, const void* target, void* workspace, size_t workspace_bytes,
                          double* loss_host);

/* ---- triplet MLP post-equaliser (no counterpart in the reference) ----
 * Widens the perturbation-triplet equaliser (src/focss/equalizer/sino.cpp,
 * term at sino.cpp:154-168) into a two-layer network over symbols.  For
 * symbol s of mode md (symbols periodic, indices mod `symbols`):
 *   T_t(s) = x_md(s+m) sum_p x_p(s+n) conj(x_p(s+m+n)),  t = (m+16) 33 + n+16,
 *   |m|, |n| <= 16 (1089 triplets); f = (Re T_0, Im T_0, Re T_1, ...) padded
 *   with zeros to FOCSS_TMLP_FEATURES;
 *   y = x + (o_0 + i o_1),  o = W2 max(W1 f + b1, 0) + b2.
 * params (device, float, focss_tmlp_param_count() values):
 *   W1 [128][FOCSS_TMLP_FEATURES] (the padding columns stay zero under
 *   training), b1 [128], W2 [2][128], b2 [2].
 * Symbols are complex64 device buffers [batch][modes][symbols], modes 1 or
 * 2, symbols a multiple of 128 in [128, 2^20], batch modes symbols <= 2^27.
 * The two long products (W1 f and its weight gradient) run on the matrix
 * cores with BF16 operands (features, W1, hidden-layer gradient) and FP32
 * accumulation.  Loss = (1/(batch modes symbols)) sum |y - target|^2.
 * Callers: post-DBP equalisation and its training. */
#define FOCSS_TMLP_HIDDEN 128
#define FOCSS_TMLP_FEATURES 2240
int focss_tmlp_param_count(void);
size_t focss_tmlp_workspace_bytes(int modes, int symbols, int batch);
/* Inference into out (must not alias in; stream-ordered), and the loss
 * against target when target is given (*loss, host double; the call then
 * synchronises). */
int focss_tmlp_forward(focss_ctx* ctx, int modes, int symbols, int batch, const float* params, const void* in,
                       void* out, const void* target, double* loss);
/* Accumulates into grad (device double [param_count], zeroed by the caller)
 * dL/dparams and into *loss (device double, may be null) the loss.
 * Stream-ordered, no host sync. */
int focss_tmlp_gradients(focss_ctx* ctx, int modes, int symbols, int batch, const float* params, const void* in,
                         const void* target, void* workspace, size_t workspace_bytes, double* grad, double* loss);
/* One training step: gradients then Adam (t = step number from 1);
 * *loss_host (may be null) receives the batch loss before the update. */
int focss_tmlp_train_step(focss_ctx* ctx, int modes, int symbols, int batch, float* params, float* adam_m,
                          float* adam_v, int t, float lr, const void* in, const void* target, void* workspace,
                          size_t workspace_bytes, double* loss_host);

/* Tuning: choose the tile shape of the two-pass kernels for one length,
 * for both passes (focss_set_tile) or for one (pass 0 = time-domain pass,
 * 1 = spectral pass).  Exposed for the autotuner; no counterpart in the
 * reference (FFTW's planner plays this role, field.cpp:530-541). */
int focss_set_tile(focss_ctx* ctx, int precision, int log2_samples, int variant);
int focss_set_tile_pass(focss_ctx* ctx, int precision, int log2_samples, int pass, int variant);
int focss_tile_variants(int precision, int log2_samples);

/* Version / capability string of the build. */
const char* focss_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* FOCSS_H_ */
