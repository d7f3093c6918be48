/*
 * focss oracle -- CPU FP64 restatement of the reference's split-step Fourier
 * hot path, written from the reference's algorithm, used ONLY as test
 * infrastructure: the parity tests, the smoke check and the bench's CPU
 * baseline leg use it as the reference results are checked against.
 * Nothing in the shipped library links it, and the GPU path never falls
 * back to it.
 *
 * Conventions restated from the reference (file:line of euav/focss):
 *   - field layout: mode-major interleaved complex<double>, element (mode,
 *     sample) at data[sample + mode*samples]           src/focss/field.cpp:140-155
 *   - forward FFT e^{-2pi i jk/N} followed by a separate x1/N pass, inverse
 *     e^{+2pi i jk/N} unscaled                           src/focss/field.cpp:343-352
 *   - omega^2 table with the Nyquist bin at +N/2        src/focss/solver/ssfm.cpp:40-49
 *   - dispersion exp(+i beta2 h/2 w^2), loss exp(-alpha h/2), Kerr
 *     exp(+i kappa gamma h P)                            src/focss/solver/ssfm.cpp:142-172
 *
 * Functions whose results no closed form, invariant or brute force pins are
 * marked "parity unpinned" in their comment.
 */
#ifndef FOCSS_ORACLE_H_
#define FOCSS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* complex<double> storage: (re, im) pairs */
typedef struct { double re, im; } orc_c128;

typedef struct {
    double alpha, beta2, gamma, length, wavelength, kappa;
} orc_fiber;                               /* src/focss/module/fiber.h:5-21 */

enum { ORC_UNIFORM = 0, ORC_LOGARITHMIC = 1, ORC_ADAPTIVE = 2 };

typedef struct {
    int grid;                 /* ORC_UNIFORM / ORC_LOGARITHMIC / ORC_ADAPTIVE */
    int total_steps;          /* UNIFORM */
    double max_phase_shift;   /* LOGARITHMIC, ADAPTIVE */
    double max_step_size;     /* LOGARITHMIC, ADAPTIVE */
} orc_grid;

/* ---- counter-based RNG (Philox4x32-10) shared bit-for-bit with the GPU ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_gauss_pair(uint64_t seed, uint32_t stream, uint64_t index, double* z0, double* z1);

/* ---- FFT: own FP64 Stockham, reference normalisation ---- */
void orc_fft_inplace(int modes, int samples, orc_c128* data);   /* forward, then x1/N */
void orc_ifft_inplace(int modes, int samples, orc_c128* data);  /* inverse, unscaled */
void orc_dft_naive(int sign, int n, const orc_c128* in, orc_c128* out); /* O(n^2), unscaled */

/* ---- Field reductions (src/focss/field.cpp:164-187) ---- */
double orc_power(int modes, int samples, const orc_c128* f, int sample);
double orc_peak_power(int modes, int samples, const orc_c128* f);
double orc_average_power(int modes, int samples, const orc_c128* f);

/* ---- SSFM (src/focss/solver/ssfm.cpp:39-172) ---- */
int orc_ssfm_run(const orc_fiber* fiber, const orc_grid* grid, double direction,
                 int modes, int samples, double dt, orc_c128* field, int* steps_done);

/* ---- Amplifier (src/focss/module/amplifier.cpp:15-42) ---- */
void orc_amplify(double gain, double noise_factor, double center_frequency, double dt,
                 uint64_t seed, uint32_t stream, int modes, int samples, orc_c128* f);
void orc_add_noise(double gain, double noise_factor, double center_frequency, double dt,
                   uint64_t seed, uint32_t stream, int modes, int samples, orc_c128* f);
void orc_drop_power(double gain, int modes, int samples, orc_c128* f);
void orc_give_power(double gain, int modes, int samples, orc_c128* f);

/* ---- link drivers (src/tasks/systems.cpp:26-91, wdm_run.cpp:26-47) ---- */
int orc_forward_propagation(const orc_fiber* fiber, int n_spans, double max_phase_shift,
                            double max_step_size, double gain, double noise_factor,
                            double center_frequency, uint64_t seed, int modes, int samples,
                            double dt, orc_c128* field, int* steps_per_span);
int orc_backward_propagation(const orc_fiber* fiber, int n_spans, int steps,
                             double max_phase_shift, double max_step_size, double gain,
                             int modes, int samples, double dt, orc_c128* field);
int orc_cd_compensation(const orc_fiber* fiber, int n_spans, int modes, int samples,
                        double dt, orc_c128* field);

/* ---- conversions (src/focss/functions.cpp:82-103) ---- */
double orc_db_to_linear(double db);
double orc_dbm_to_watts(double dbm);
double orc_db_to_natural(double db);
double orc_disp_to_beta2(double dispersion, double wavelength);

/* ---- modulation / DSP harness (src/focss/processing/modulation.cpp) ---- */
void orc_gray16(int code, double* re, double* im);
void orc_qam16_symbols(uint64_t seed, int n, orc_c128* out, int* codes);
void orc_rrc_pulse(int n_symbols, int osf, double roll_off, orc_c128* out);
void orc_filter_inplace(int modes, int samples, orc_c128* f, int taps, const orc_c128* h);
void orc_rrc_shaping(int modes, int n_symbols, const orc_c128* symbols, int osf,
                     double roll_off, double gain, orc_c128* out);
void orc_rrc_sampling(int modes, int samples, const orc_c128* signal, int osf,
                      double roll_off, double gain, orc_c128* out);
void orc_decimate(int modes, int samples, const orc_c128* in, int factor, orc_c128* out);
void orc_frequency_shift(int modes, int samples, double dt, double frequency, orc_c128* f);
void orc_bandpass_filter(int modes, int samples, double dt, double frequency, orc_c128* f);
void orc_wdm_select(int modes, int samples, double dt, const orc_c128* in, int channel_index,
                    double channel_bandwidth, orc_c128* out);
unsigned orc_demodulate_16qam(double re, double im);
void orc_hd_sequence_16qam(int modes, int samples, const orc_c128* f, int* codes);
double orc_ber_16qam(int modes, int samples, const orc_c128* tx, const orc_c128* rx, int cut);
double orc_q2_factor(int modes, int samples, const orc_c128* tx, const orc_c128* rx, int cut);
double orc_pse_estimate(int size, const orc_c128* desired, const orc_c128* actual, int angle_steps);
orc_c128 orc_scalar_estimate(int size, const orc_c128* desired, const orc_c128* actual);
double orc_homothety_estimate(int size, const orc_c128* desired, const orc_c128* actual);

/* ---- perturbation-triplet equalizer (src/focss/equalizer/sino.cpp) ---- */
void orc_sino_term(int modes, int samples, const orc_c128* f, int mode, int k, int m, int n,
                   orc_c128* out);
void orc_sino_cubic(int radius, const orc_c128* weights, int modes, int samples,
                    const orc_c128* in, orc_c128* out);
int orc_sino_n_features(int radius, int type);
void orc_sino_features(int radius, int type, orc_c128 linear_weight, int modes, int samples,
                       const orc_c128* tx, const orc_c128* rx, orc_c128* features,
                       orc_c128* responses);
int orc_sino_train(int radius, int type, orc_c128 linear_weight_in, int modes, int samples,
                   const orc_c128* tx, const orc_c128* rx, orc_c128* weights,
                   orc_c128* intercept, orc_c128* linear_weight);
int orc_complex_lstsq(int rows, int cols, const orc_c128* a_colmajor, const orc_c128* b,
                      orc_c128* x);

/* ---- WDM dual-pol triplet equalizer (src/focss/equalizer/wdm_sino.cpp) ---- */
void orc_wdmppe_spm(int c_radius, const orc_c128* C, int samples, const orc_c128* A6,
                    orc_c128* out2);
void orc_wdmppe_xpm(int d_radius, const orc_c128* D, int samples, const orc_c128* A6,
                    orc_c128* out2);
int orc_wdmppe_n_features(int c_radius, int d_radius);
void orc_wdmppe_features(int c_radius, int d_radius, int samples, const orc_c128* tx6,
                         const orc_c128* rx6, orc_c128* features, orc_c128* responses);
int orc_wdmppe_train(int c_radius, int d_radius, int samples, const orc_c128* tx6, const orc_c128* rx6,
                     orc_c128* C, orc_c128* D, orc_c128* intercept, orc_c128* linear_weight);

/* ---- learned few-step DBP (no reference counterpart; measured against plain DBP) ----
 * fir [steps][taps] complex, kerr [steps][ktaps] real intensity taps of the
 * nonlinear phase (ktaps = 1: exp(i theta |y|^2)). */
double orc_ldbp_forward(int steps, int taps, int ktaps, int modes, const orc_c128* fir, const double* kerr,
                        int samples, const orc_c128* in, orc_c128* out,
                        const orc_c128* target);
double orc_ldbp_backward(int steps, int taps, int ktaps, int modes, const orc_c128* fir, const double* kerr,
                         int samples, const orc_c128* in, const orc_c128* target,
                         orc_c128* grad_fir, double* grad_kerr);

/* ---- triplet MLP post-equaliser (no reference counterpart; its triplet is
 * the term of src/focss/equalizer/sino.cpp:154-168, with periodic indices,
 * over the full square |m|, |n| <= 16) ----
 * params: W1 [128][2240] (features 2t, 2t+1 = Re, Im of triplet
 * t = (m+16) 33 + n+16; columns from 2178 on unused), b1 [128], W2 [2][128],
 * b2 [2].  Symbols [batch][modes][symbols].  bf16 != 0 rounds the features
 * and W1 (and, in the backward, the hidden-layer gradient before the W1
 * product) to bfloat16, round to nearest even, as the matrix-core kernels
 * do; everything else is FP64. */
int orc_tmlp_n_params(void);
double orc_tmlp_forward(int modes, int symbols, int batch, const double* params, int bf16,
                        const orc_c128* in, orc_c128* out, const orc_c128* target);
double orc_tmlp_backward(int modes, int symbols, int batch, const double* params, int bf16,
                         const orc_c128* in, const orc_c128* target, double* grad);

#ifdef __cplusplus
}
#endif
#endif
