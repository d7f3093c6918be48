/*
 * focss oracle -- CPU FP64 restatement of the reference's SSFM / DBP /
 * triplet-equalizer hot path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke
 * check and the bench's CPU-baseline leg use it as the reference results are
 * checked against; the shipped
 * library never links or calls it.
 *
 * Every function cites the reference lines (euav/focss) it restates.  The
 * FFT is this file's own Stockham transform (the reference uses FFTW, which
 * is not available here); its results are pinned by a brute-force DFT and by
 * numpy.fft in the tests.
 */
#include "focss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.14159265358979323846264338;      /* definitions.h:14 */
static const double kPlanck = 6.62607004081818181818181e-34; /* definitions.h:15 */
static const double kLightSpeed = 299792458.0;               /* definitions.h:16 */

/* ------------------------------------------------------------------ */
/* complex helpers                                                     */
/* ------------------------------------------------------------------ */
static inline orc_c128 c_mul(orc_c128 a, orc_c128 b) {
    orc_c128 r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline orc_c128 c_add(orc_c128 a, orc_c128 b) {
    orc_c128 r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline orc_c128 c_sub(orc_c128 a, orc_c128 b) {
    orc_c128 r = {a.re - b.re, a.im - b.im};
    return r;
}
static inline orc_c128 c_conj(orc_c128 a) {
    orc_c128 r = {a.re, -a.im};
    return r;
}
static inline orc_c128 c_scale(orc_c128 a, double s) {
    orc_c128 r = {a.re * s, a.im * s};
    return r;
}
static inline double c_norm(orc_c128 a) { return a.re * a.re + a.im * a.im; }
/* i_exp(x) = (cos x, sin x)                          functions.cpp:103 */
static inline orc_c128 i_exp(double x) {
    orc_c128 r = {cos(x), sin(x)};
    return r;
}
static inline orc_c128 c_div(orc_c128 a, orc_c128 b) {
    /* std::complex operator/ semantics are not pinned bit-for-bit; this is
       the textbook formula */
    double d = b.re * b.re + b.im * b.im;
    orc_c128 r = {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
    return r;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11).  Same algorithm runs on the GPU.        */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words (a>>5, b>>6 composition). */
static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Box-Muller pair from one Philox block.  counter = (index lo, index hi,
   stream, 0x0A5E0A5E), key = seed.  Replaces std::normal_distribution
   (amplifier.cpp:25), whose realisation is implementation-defined. */
void orc_gauss_pair(uint64_t seed, uint32_t stream, uint64_t index, double* z0, double* z1) {
    uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x0A5E0A5Eu};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    double u1 = 1.0 - u53(o[0], o[1]); /* (0,1] */
    double u2 = u53(o[2], o[3]);
    double r = sqrt(-2.0 * log(u1));
    double t = 2.0 * kPi * u2;
    *z0 = r * cos(t);
    *z1 = r * sin(t);
}

/* ------------------------------------------------------------------ */
/* FFT of any length N (FFTW plans any N: field.cpp:530-541).           */
/* Stockham auto-sort (Govindaraju et al., "High performance discrete   */
/* Fourier transforms on graphics processors", SC'08, Fig. 2): one      */
/* stage per factor of N, radix 4 while 4 divides what is left, then 2, */
/* then the odd primes, each a direct length-r DFT.  A prime factor     */
/* above 64 goes through Bluestein's chirp-z identity (Bluestein, 1970; */
/* jk = (j^2 + k^2 - (k-j)^2)/2) as a circular convolution of a         */
/* power-of-two length M >= 2N-1.                                        */
/* The twiddle table W_N^m = (cos 2 pi m/N, sign sin 2 pi m/N) is kept   */
/* per thread for the last few (N, sign), so repeated transforms of one */
/* length do not recompute it; its values do not depend on the cache.    */
/* ------------------------------------------------------------------ */
#define ORC_TW_SLOTS 4
typedef struct {
    int n, sign;
    orc_c128* w;
} tw_slot;
static __thread tw_slot tw_cache[ORC_TW_SLOTS];
static __thread int tw_next;

static const orc_c128* twiddles(int n, int sign) {
    for (int i = 0; i < ORC_TW_SLOTS; ++i)
        if (tw_cache[i].w && tw_cache[i].n == n && tw_cache[i].sign == sign) return tw_cache[i].w;
    tw_slot* s = &tw_cache[tw_next];
    tw_next = (tw_next + 1) % ORC_TW_SLOTS;
    free(s->w);
    s->w = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)n);
    s->n = n;
    s->sign = sign;
    for (int m = 0; m < n; ++m) {
        double a = 2.0 * kPi * (double)m / (double)n;
        s->w[m].re = cos(a);
        s->w[m].im = sign * sin(a);
    }
    return s->w;
}

/* smallest factor of n to take next: 4, 2, then odd primes; 0 if the
   remaining n is a prime above 64 */
static int next_radix(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p <= 64 && p <= n; p += 2)
        if (n % p == 0) return p;
    return n == 1 ? 1 : 0;
}

static void fft_any(int n, int sign, orc_c128* x);

static void stockham(int n, int sign, orc_c128* x) {
    if (n <= 1) return;
    const orc_c128* w = twiddles(n, sign);
    orc_c128* work = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)n);
    orc_c128* src = x;
    orc_c128* dst = work;
    orc_c128 v[64], y[64];
    for (int ns = 1; ns < n;) {
        int radix = next_radix(n / ns);
        int stride = n / radix;
        int tw_step = n / (ns * radix);
        for (int j = 0; j < stride; ++j) {
            int k = j % ns;
            for (int r = 0; r < radix; ++r) {
                orc_c128 a = src[j + r * stride];
                v[r] = (r == 0) ? a : c_mul(a, w[(long)r * k * tw_step]);
            }
            if (radix == 2) {
                y[0] = c_add(v[0], v[1]);
                y[1] = c_sub(v[0], v[1]);
            } else if (radix == 4) {
                /* W4 = exp(sign * 2 pi i / 4) = sign * i */
                orc_c128 a0 = c_add(v[0], v[2]);
                orc_c128 a1 = c_sub(v[0], v[2]);
                orc_c128 b0 = c_add(v[1], v[3]);
                orc_c128 b1 = c_sub(v[1], v[3]);
                orc_c128 jb1 = {-sign * b1.im, sign * b1.re}; /* (sign i) * b1 */
                y[0] = c_add(a0, b0);
                y[2] = c_sub(a0, b0);
                y[1] = c_add(a1, jb1);
                y[3] = c_sub(a1, jb1);
            } else {
                /* direct length-r DFT, W_r^{pq} = W_N^{(pq mod r) N/r} */
                for (int p = 0; p < radix; ++p) {
                    orc_c128 s = v[0];
                    for (int q = 1; q < radix; ++q) s = c_add(s, c_mul(v[q], w[(long)((p * q) % radix) * stride]));
                    y[p] = s;
                }
            }
            int base = (j / ns) * ns * radix + k;
            for (int r = 0; r < radix; ++r) dst[base + r * ns] = y[r];
        }
        orc_c128* t = src;
        src = dst;
        dst = t;
        ns *= radix;
    }
    if (src != x) memcpy(x, src, sizeof(orc_c128) * (size_t)n);
    free(work);
}

/* Bluestein: X_k = c_k sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(sign pi i j^2/N) */
static void bluestein(int n, int sign, orc_c128* x) {
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    orc_c128* chirp = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)n);
    orc_c128* a = (orc_c128*)calloc((size_t)m, sizeof(orc_c128));
    orc_c128* b = (orc_c128*)calloc((size_t)m, sizeof(orc_c128));
    for (int j = 0; j < n; ++j) {
        long long q = ((long long)j * j) % (2LL * n); /* j^2 mod 2N keeps the angle small */
        double ang = kPi * (double)q / (double)n;
        chirp[j].re = cos(ang);
        chirp[j].im = sign * sin(ang);
    }
    for (int j = 0; j < n; ++j) a[j] = c_mul(x[j], chirp[j]);
    b[0] = c_conj(chirp[0]);
    for (int j = 1; j < n; ++j) b[j] = b[m - j] = c_conj(chirp[j]);
    stockham(m, -1, a);
    stockham(m, -1, b);
    for (int i = 0; i < m; ++i) a[i] = c_mul(a[i], b[i]);
    stockham(m, +1, a);
    double inv = 1.0 / (double)m;
    for (int k = 0; k < n; ++k) x[k] = c_scale(c_mul(a[k], chirp[k]), inv);
    free(b);
    free(a);
    free(chirp);
}

/* unscaled DFT of length n, exponent sign `sign` */
static void fft_any(int n, int sign, orc_c128* x) {
    if (n <= 1) return;
    int r = n;
    while (r > 1) {
        int p = next_radix(r);
        if (p == 0) break;
        r /= p;
    }
    if (r == 1)
        stockham(n, sign, x);
    else
        bluestein(n, sign, x);
}

/* Field::fft_inplace: FFTW_FORWARD then x1/N        field.cpp:343-347 */
void orc_fft_inplace(int modes, int samples, orc_c128* data) {
    for (int m = 0; m < modes; ++m) fft_any(samples, -1, data + (size_t)m * samples);
    double s = 1.0 / (double)samples;
    for (long i = 0; i < (long)modes * samples; ++i) data[i] = c_scale(data[i], s);
}

/* Field::ifft_inplace: FFTW_BACKWARD, unscaled      field.cpp:349-352 */
void orc_ifft_inplace(int modes, int samples, orc_c128* data) {
    for (int m = 0; m < modes; ++m) fft_any(samples, +1, data + (size_t)m * samples);
}

void orc_dft_naive(int sign, int n, const orc_c128* in, orc_c128* out) {
    for (int k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        for (int j = 0; j < n; ++j) {
            long double a = 2.0L * 3.141592653589793238462643383279L * (long double)(((long)j * k) % n) / n;
            long double c = cosl(a), s = sign * sinl(a);
            re += in[j].re * c - in[j].im * s;
            im += in[j].re * s + in[j].im * c;
        }
        out[k].re = (double)re;
        out[k].im = (double)im;
    }
}

/* ------------------------------------------------------------------ */
/* Field power reductions                              field.cpp:164-187 */
/* ------------------------------------------------------------------ */
double orc_power(int modes, int samples, const orc_c128* f, int sample) {
    double p = 0;
    for (long i = sample; i < (long)modes * samples; i += samples) p += c_norm(f[i]);
    return p;
}

double orc_peak_power(int modes, int samples, const orc_c128* f) {
    double peak = 0;
    for (int s = 0; s < samples; ++s) {
        double p = orc_power(modes, samples, f, s);
        if (peak < p) peak = p;
    }
    return peak;
}

double orc_average_power(int modes, int samples, const orc_c128* f) {
    double sum = 0;
    for (long i = 0; i < (long)modes * samples; ++i) sum += c_norm(f[i]);
    return sum / samples;
}

static void scale_field(long size, orc_c128* f, double s) {
    for (long i = 0; i < size; ++i) f[i] = c_scale(f[i], s);
}

/* ------------------------------------------------------------------ */
/* SSFM                                                ssfm.cpp:39-172  */
/* ------------------------------------------------------------------ */
typedef struct {
    orc_fiber fiber;
    orc_grid grid;
    int modes, samples;
    double* disp; /* omega^2, FFT order */
} ssfm_ctx;

/* linear_step: x e^{-alpha h/2}; fft; x exp(i beta2 h/2 w^2); ifft    ssfm.cpp:142-150 */
static void linear_step(const ssfm_ctx* c, orc_c128* f, double step) {
    long size = (long)c->modes * c->samples;
    scale_field(size, f, exp(-0.5 * c->fiber.alpha * step));
    orc_fft_inplace(c->modes, c->samples, f);
    double arg = 0.5 * c->fiber.beta2 * step;
    for (long i = 0; i < size; ++i) f[i] = c_mul(f[i], i_exp(arg * c->disp[i % c->samples]));
    orc_ifft_inplace(c->modes, c->samples, f);
}

/* nofft_linear_step                                   ssfm.cpp:152-158 */
static void nofft_linear_step(const ssfm_ctx* c, orc_c128* f, double step) {
    long size = (long)c->modes * c->samples;
    scale_field(size, f, exp(-0.5 * c->fiber.alpha * step));
    double arg = 0.5 * c->fiber.beta2 * step;
    for (long i = 0; i < size; ++i) f[i] = c_mul(f[i], i_exp(arg * c->disp[i % c->samples]));
}

/* nonlinear_step: power is recomputed per element from the partially
   rotated field, exactly as the reference loop does   ssfm.cpp:160-164 */
static void nonlinear_step(const ssfm_ctx* c, orc_c128* f, double step) {
    long size = (long)c->modes * c->samples;
    double arg = c->fiber.kappa * c->fiber.gamma * step;
    for (long i = 0; i < size; ++i) {
        double p = orc_power(c->modes, c->samples, f, (int)(i % c->samples));
        f[i] = c_mul(f[i], i_exp(arg * p));
    }
}

/* fft_nonlinear_step                                  ssfm.cpp:166-172 */
static void fft_nonlinear_step(const ssfm_ctx* c, orc_c128* f, double step) {
    orc_ifft_inplace(c->modes, c->samples, f);
    nonlinear_step(c, f, step);
    orc_fft_inplace(c->modes, c->samples, f);
}

/* uniform_solve                                       ssfm.cpp:62-72 */
static void uniform_solve(const ssfm_ctx* c, orc_c128* f, double direction) {
    double step = c->fiber.length / c->grid.total_steps;
    step = direction * step;
    linear_step(c, f, 0.5 * step);
    for (int i = 0; i < c->grid.total_steps; ++i) {
        nonlinear_step(c, f, step);
        linear_step(c, f, step);
    }
    linear_step(c, f, -0.5 * step);
}

/* estimate_logarithmic_steps                          ssfm.cpp:120-134 */
static int estimate_logarithmic_steps(const ssfm_ctx* c, const orc_c128* f) {
    double alpha = c->fiber.alpha;
    double l_nl = 1.0 / c->fiber.gamma / orc_average_power(c->modes, c->samples, f);
    double cl_nl_kappa = c->grid.max_phase_shift * l_nl / c->fiber.kappa;
    double l_eff = (1.0 - exp(-alpha * c->fiber.length)) / alpha;
    double l_nl_eff = (1.0 - exp(-alpha * cl_nl_kappa)) / alpha;
    int estimated = (int)(l_eff / l_nl_eff);
    int minimum = (int)(c->fiber.length / c->grid.max_step_size);
    return minimum > estimated ? minimum : estimated;
}

/* logarithmic_solve                                   ssfm.cpp:74-90 */
static int logarithmic_solve(const ssfm_ctx* c, orc_c128* f, double direction) {
    int n = estimate_logarithmic_steps(c, f);
    double inv_alpha = -1.0 / c->fiber.alpha * direction;
    double inv_gain = exp(-c->fiber.alpha * c->fiber.length) - 1;
    inv_gain /= n;
    orc_fft_inplace(c->modes, c->samples, f);
    for (int i = 0; i < n; ++i) {
        double step = inv_alpha * log(1 + inv_gain / (1 + inv_gain * i));
        nofft_linear_step(c, f, 0.5 * step);
        fft_nonlinear_step(c, f, step);
        nofft_linear_step(c, f, 0.5 * step);
    }
    orc_ifft_inplace(c->modes, c->samples, f);
    return n;
}

/* estimate_next_adaptive_step                         ssfm.cpp:136-140 */
static double estimate_next_adaptive_step(const ssfm_ctx* c, const orc_c128* f) {
    double argument = fabs(c->fiber.kappa * c->fiber.gamma);
    double phase_shift = argument * orc_peak_power(c->modes, c->samples, f);
    double h = c->grid.max_phase_shift / phase_shift;
    /* std::min(a, b) returns a unless b < a */
    return (h < c->grid.max_step_size) ? h : c->grid.max_step_size;
}

/* adaptive_solve                                      ssfm.cpp:92-118 */
static int adaptive_solve(const ssfm_ctx* c, orc_c128* f, double direction) {
    double distance = 0;
    double next_step = estimate_next_adaptive_step(c, f);
    double step = next_step;
    int steps_done = 1;
    orc_fft_inplace(c->modes, c->samples, f);
    while (distance + step < c->fiber.length) {
        nofft_linear_step(c, f, 0.5 * step * direction);
        orc_ifft_inplace(c->modes, c->samples, f);
        nonlinear_step(c, f, step * direction);
        next_step = estimate_next_adaptive_step(c, f);
        orc_fft_inplace(c->modes, c->samples, f);
        nofft_linear_step(c, f, 0.5 * step * direction);
        distance += step;
        step = next_step;
        steps_done++;
    }
    step = c->fiber.length - distance;
    nofft_linear_step(c, f, 0.5 * step * direction);
    fft_nonlinear_step(c, f, step * direction);
    nofft_linear_step(c, f, 0.5 * step * direction);
    orc_ifft_inplace(c->modes, c->samples, f);
    return steps_done;
}

/* SSFM::run: omega^2 table in FFT order, then the grid   ssfm.cpp:39-60 */
int orc_ssfm_run(const orc_fiber* fiber, const orc_grid* grid, double direction, int modes,
                 int samples, double dt, orc_c128* field, int* steps_done) {
    if (modes <= 0 || samples <= 0 || dt <= 0) return -1;
    if (grid->grid == ORC_UNIFORM && grid->total_steps <= 0) return -2;
    if (grid->grid == ORC_ADAPTIVE && !(grid->max_step_size > 0)) return -3;
    if (grid->grid == ORC_LOGARITHMIC && fiber->alpha == 0) return -4;
    ssfm_ctx c;
    c.fiber = *fiber;
    c.grid = *grid;
    c.modes = modes;
    c.samples = samples;
    c.disp = (double*)malloc(sizeof(double) * (size_t)samples);
    double dw = 2.0 * kPi / (dt * samples); /* Field::dw  field.cpp:220 */
    for (int i = 0; i <= samples / 2; ++i) c.disp[i] = (double)i * dw;
    for (int i = samples / 2 + 1; i < samples; ++i) c.disp[i] = (double)(i - samples) * dw;
    for (int i = 0; i < samples; ++i) c.disp[i] *= c.disp[i];
    int n = 0;
    if (grid->grid == ORC_UNIFORM) {
        uniform_solve(&c, field, direction);
        n = grid->total_steps;
    } else if (grid->grid == ORC_LOGARITHMIC) {
        n = logarithmic_solve(&c, field, direction);
    } else {
        n = adaptive_solve(&c, field, direction);
    }
    free(c.disp);
    if (steps_done) *steps_done = n;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Amplifier                                    amplifier.cpp:15-42     */
/* ------------------------------------------------------------------ */
void orc_add_noise(double gain, double noise_factor, double center_frequency, double dt,
                   uint64_t seed, uint32_t stream, int modes, int samples, orc_c128* f) {
    double variance = kPlanck * center_frequency; /* amplifier.cpp:21-23 */
    variance *= noise_factor * (gain - 1) / 2;
    variance *= 1.0 / dt; /* Field::bandwidth */
    double sigma = sqrt(variance / 2);
    for (long i = 0; i < (long)modes * samples; ++i) {
        double z0, z1;
        orc_gauss_pair(seed, stream, (uint64_t)i, &z0, &z1);
        f[i].re += sigma * z0;
        f[i].im += sigma * z1;
    }
}

void orc_amplify(double gain, double noise_factor, double center_frequency, double dt,
                 uint64_t seed, uint32_t stream, int modes, int samples, orc_c128* f) {
    scale_field((long)modes * samples, f, sqrt(gain));
    if (noise_factor != 0)
        orc_add_noise(gain, noise_factor, center_frequency, dt, seed, stream, modes, samples, f);
}

void orc_drop_power(double gain, int modes, int samples, orc_c128* f) {
    if (gain < 1)
        scale_field((long)modes * samples, f, sqrt(gain));
    else
        scale_field((long)modes * samples, f, sqrt(1 / gain));
}

void orc_give_power(double gain, int modes, int samples, orc_c128* f) {
    if (gain > 1)
        scale_field((long)modes * samples, f, sqrt(gain));
    else
        scale_field((long)modes * samples, f, sqrt(1 / gain));
}

/* ------------------------------------------------------------------ */
/* Link drivers                               systems.cpp:26-91         */
/* ------------------------------------------------------------------ */
/* forward_propagation: per span ADAPTIVE SSFM then amplify (ASE stream =
   span index)                                   systems.cpp:26-43 */
int orc_forward_propagation(const orc_fiber* fiber, int n_spans, double max_phase_shift,
                            double max_step_size, double gain, double noise_factor,
                            double center_frequency, uint64_t seed, int modes, int samples,
                            double dt, orc_c128* field, int* steps_per_span) {
    orc_grid g = {ORC_ADAPTIVE, 0, max_phase_shift, max_step_size};
    for (int s = 0; s < n_spans; ++s) {
        int n = 0;
        int rc = orc_ssfm_run(fiber, &g, 1.0, modes, samples, dt, field, &n);
        if (rc) return rc;
        if (steps_per_span) steps_per_span[s] = n;
        orc_amplify(gain, noise_factor, center_frequency, dt, seed, (uint32_t)s, modes, samples,
                    field);
    }
    return 0;
}

/* backward_propagation (DBP): per span drop_power then BACKWARD SSFM,
   UNIFORM(steps) or LOGARITHMIC when steps == 0    systems.cpp:73-91 */
int orc_backward_propagation(const orc_fiber* fiber, int n_spans, int steps,
                             double max_phase_shift, double max_step_size, double gain,
                             int modes, int samples, double dt, orc_c128* field) {
    orc_grid g;
    if (steps == 0) {
        g.grid = ORC_LOGARITHMIC;
        g.total_steps = 0;
    } else {
        g.grid = ORC_UNIFORM;
        g.total_steps = steps;
    }
    g.max_phase_shift = max_phase_shift;
    g.max_step_size = max_step_size;
    for (int s = 0; s < n_spans; ++s) {
        orc_drop_power(gain, modes, samples, field);
        int rc = orc_ssfm_run(fiber, &g, -1.0, modes, samples, dt, field, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* cd_compensation: SSFM(fiber.dispersive() * n_spans, 1) BACKWARD
                                                  systems.cpp:63-71, fiber.cpp:16-45 */
int orc_cd_compensation(const orc_fiber* fiber, int n_spans, int modes, int samples, double dt,
                        orc_c128* field) {
    orc_fiber d = *fiber;
    d.alpha = 0;
    d.gamma = 0;
    d.length *= n_spans;
    orc_grid g = {ORC_UNIFORM, 1, 0, 0};
    return orc_ssfm_run(&d, &g, -1.0, modes, samples, dt, field, NULL);
}

/* ------------------------------------------------------------------ */
/* Conversions                                 functions.cpp:82-96      */
/* ------------------------------------------------------------------ */
double orc_db_to_linear(double db) { return pow(10, db / 10); }
double orc_dbm_to_watts(double dbm) { return 1e-3 * pow(10, dbm / 10); }
double orc_db_to_natural(double db) { return db * log(10) / 10; }
double orc_disp_to_beta2(double dispersion, double wavelength) {
    return -wavelength * wavelength * dispersion / (2 * kPi * kLightSpeed);
}

/* ------------------------------------------------------------------ */
/* Modulation / DSP harness                    modulation.cpp           */
/* ------------------------------------------------------------------ */
/* Gray 16QAM table                             modulation.cpp:9-17 */
static const signed char kGray[16][2] = {{-3, -3}, {-1, -3}, {3, -3}, {1, -3},
                                         {-3, -1}, {-1, -1}, {3, -1}, {1, -1},
                                         {-3, 3},  {-1, 3},  {3, 3},  {1, 3},
                                         {-3, 1},  {-1, 1},  {3, 1},  {1, 1}};

void orc_gray16(int code, double* re, double* im) {
    *re = kGray[code][0] / sqrt(10.0);
    *im = kGray[code][1] / sqrt(10.0);
}

/* i.i.d. uniform symbols from Philox (replaces the implementation-defined
   std::shuffle of modulation.cpp:76-83) */
void orc_qam16_symbols(uint64_t seed, int n, orc_c128* out, int* codes) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int i = 0; i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, 0u, 0u, 0x16A016A0u};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        int code = (int)(o[0] & 15u);
        if (codes) codes[i] = code;
        orc_gray16(code, &out[i].re, &out[i].im);
    }
}

/* pulse::rrc                                    modulation.cpp:47-74 */
void orc_rrc_pulse(int n_symbols, int osf, double roll_off, orc_c128* out) {
    int size = n_symbols * osf;
    for (int i = 0; i < size; ++i) {
        double t = (double)i / osf - n_symbols / 2.0;
        double v = 0;
        if (t == 0) {
            v = 1 - roll_off + 4 * roll_off / kPi;
        } else if (fabs(t) == 1 / roll_off / 4) {
            v += (1 + 2 / kPi) * sin(kPi / roll_off / 4);
            v += (1 - 2 / kPi) * cos(kPi / roll_off / 4);
            v *= roll_off / sqrt(2.0);
        } else {
            v += cos(kPi * t * (1 + roll_off));
            v *= 4 * roll_off * t;
            v += sin(kPi * t * (1 - roll_off));
            v /= 1 - 16 * roll_off * roll_off * t * t;
            v /= kPi * t;
        }
        out[i].re = v;
        out[i].im = 0;
    }
    double power = 0;
    for (int i = 0; i < size; ++i) power += c_norm(out[i]);
    for (int i = 0; i < size; ++i) out[i] = c_scale(out[i], size / power);
}

/* Field::filter_inplace: circular FFT filter with the impulse response
   centred at sample 0 (odd lengths drop the last tap)   field.cpp:358-372 */
void orc_filter_inplace(int modes, int samples, orc_c128* f, int taps, const orc_c128* h) {
    orc_c128* zp = (orc_c128*)calloc((size_t)samples, sizeof(orc_c128));
    int half = taps / 2;
    int shift = samples - half;
    for (int i = 0; i < taps / 2; ++i) {
        zp[i] = h[i + half];
        zp[shift + i] = h[i];
    }
    orc_fft_inplace(modes, samples, f);
    orc_fft_inplace(1, samples, zp); /* focss::fft(ComplexVector) also scales by 1/N */
    for (int m = 0; m < modes; ++m)
        for (int s = 0; s < samples; ++s)
            f[(long)m * samples + s] = c_mul(f[(long)m * samples + s], zp[s]);
    orc_ifft_inplace(modes, samples, f);
    free(zp);
}

/* rrc_shaping: upsample(osf), RRC filter, x sqrt(gain)*osf   modulation.cpp:86-90 */
void orc_rrc_shaping(int modes, int n_symbols, const orc_c128* symbols, int osf, double roll_off,
                     double gain, orc_c128* out) {
    int samples = n_symbols * osf;
    orc_c128* rrc = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)samples);
    orc_rrc_pulse(n_symbols, osf, roll_off, rrc);
    memset(out, 0, sizeof(orc_c128) * (size_t)modes * samples);
    for (long i = 0; i < (long)modes * n_symbols; ++i) out[(long)osf * i] = symbols[i];
    orc_filter_inplace(modes, samples, out, samples, rrc);
    scale_field((long)modes * samples, out, sqrt(gain) * osf);
    free(rrc);
}

/* rrc_sampling: RRC filter then downsample(osf), x sqrt(gain)  modulation.cpp:92-96 */
void orc_rrc_sampling(int modes, int samples, const orc_c128* signal, int osf, double roll_off,
                      double gain, orc_c128* out) {
    int n_symbols = samples / osf;
    orc_c128* rrc = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)samples);
    orc_c128* tmp = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)modes * samples);
    orc_rrc_pulse(n_symbols, osf, roll_off, rrc);
    memcpy(tmp, signal, sizeof(orc_c128) * (size_t)modes * samples);
    orc_filter_inplace(modes, samples, tmp, samples, rrc);
    /* Field::downsample on a mode-major buffer takes every osf-th element of
       the flat array (field.cpp:285-297) */
    for (long i = 0; i < (long)modes * n_symbols; ++i)
        out[i] = c_scale(tmp[i * osf], sqrt(gain));
    free(tmp);
    free(rrc);
}

/* Field::decimate: spectral truncation            field.cpp:299-320 */
void orc_decimate(int modes, int samples, const orc_c128* in, int factor, orc_c128* out) {
    int ds = samples / factor;
    orc_c128* spec = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)modes * samples);
    memcpy(spec, in, sizeof(orc_c128) * (size_t)modes * samples);
    orc_fft_inplace(modes, samples, spec);
    int cutoff = samples / factor / 2;
    int shift = samples - ds;
    for (int m = 0; m < modes; ++m) {
        for (int s = 0; s < cutoff; ++s) out[(long)m * ds + s] = spec[(long)m * samples + s];
        for (int s = cutoff; s < ds; ++s) out[(long)m * ds + s] = spec[(long)m * samples + s + shift];
    }
    orc_ifft_inplace(modes, ds, out);
    free(spec);
}

/* frequency_shift                                 modulation.cpp:255-264 */
void orc_frequency_shift(int modes, int samples, double dt, double frequency, orc_c128* f) {
    double wdt = 2 * kPi * frequency * dt;
    for (int m = 0; m < modes; ++m)
        for (int s = 0; s < samples; ++s)
            f[(long)m * samples + s] = c_mul(f[(long)m * samples + s], i_exp(wdt * s));
}

/* bandpass_filter                                 modulation.cpp:266-277 */
void orc_bandpass_filter(int modes, int samples, double dt, double frequency, orc_c128* f) {
    int cut = (int)(frequency * (samples * dt)) + 1;
    orc_fft_inplace(modes, samples, f);
    for (int m = 0; m < modes; ++m)
        for (int s = cut; s <= samples - cut; ++s) {
            f[(long)m * samples + s].re = 0;
            f[(long)m * samples + s].im = 0;
        }
    orc_ifft_inplace(modes, samples, f);
}

/* wdm_select                                      modulation.cpp:248-253 */
void orc_wdm_select(int modes, int samples, double dt, const orc_c128* in, int channel_index,
                    double channel_bandwidth, orc_c128* out) {
    memcpy(out, in, sizeof(orc_c128) * (size_t)modes * samples);
    orc_frequency_shift(modes, samples, dt, -channel_bandwidth * channel_index, out);
    orc_bandpass_filter(modes, samples, dt, 0.5 * channel_bandwidth, out);
}

/* demodulate_16qam: nearest point, strict '>' so ties keep the lower code
                                                   modulation.cpp:98-110 */
unsigned orc_demodulate_16qam(double re, double im) {
    unsigned argmin = 0;
    double gr, gi;
    orc_gray16(0, &gr, &gi);
    double mn = (re - gr) * (re - gr) + (im - gi) * (im - gi);
    for (unsigned code = 1; code < 16; ++code) {
        orc_gray16((int)code, &gr, &gi);
        double d = (re - gr) * (re - gr) + (im - gi) * (im - gi);
        if (mn > d) {
            mn = d;
            argmin = code;
        }
    }
    return argmin;
}

/* hd_sequence_16qam                               modulation.cpp:148-156 */
void orc_hd_sequence_16qam(int modes, int samples, const orc_c128* f, int* codes) {
    double scale = sqrt(modes / orc_average_power(modes, samples, f));
    for (long i = 0; i < (long)modes * samples; ++i)
        codes[i] = (int)orc_demodulate_16qam(scale * f[i].re, scale * f[i].im);
}

static unsigned bit_dist(unsigned a, unsigned b) {
    unsigned d = a ^ b, n = 0;
    while (d) {
        d &= d - 1;
        n++;
    }
    return n;
}

/* bit_error_rate_16qam (cut = 0 is the no-cut overload)   modulation.cpp:170-202 */
double orc_ber_16qam(int modes, int samples, const orc_c128* tx, const orc_c128* rx, int cut) {
    double scale = sqrt(modes / orc_average_power(modes, samples, tx));
    unsigned long errors = 0;
    for (int m = 0; m < modes; ++m)
        for (int s = cut; s < samples - cut; ++s) {
            long i = (long)m * samples + s;
            unsigned a = orc_demodulate_16qam(scale * tx[i].re, scale * tx[i].im);
            unsigned b = orc_demodulate_16qam(scale * rx[i].re, scale * rx[i].im);
            errors += bit_dist(a, b);
        }
    return (double)errors / (double)(4L * modes * samples - 8L * cut);
}

/* evm2_factor / q2_factor                         functions.cpp:27-80 */
double orc_q2_factor(int modes, int samples, const orc_c128* tx, const orc_c128* rx, int cut) {
    double num = 0, den = 0;
    for (int m = 0; m < modes; ++m)
        for (int s = cut; s < samples - cut; ++s) {
            long i = (long)m * samples + s;
            num += c_norm(c_sub(tx[i], rx[i]));
            den += c_norm(tx[i]);
        }
    return -10 * log10(num / den);
}

/* PhaseShiftEqualizer::estimate: brute-force angle search
                                                   phase_shift_equalizer.cpp:30-46 */
double orc_pse_estimate(int size, const orc_c128* desired, const orc_c128* actual,
                        int angle_steps) {
    double best = orc_q2_factor(1, size, desired, actual, 0);
    double argmax = 0;
    orc_c128* rot = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)size);
    for (int i = 0; i < angle_steps; ++i) {
        double angle = 2 * kPi * (double)i / angle_steps;
        orc_c128 r = i_exp(angle);
        for (int k = 0; k < size; ++k) rot[k] = c_mul(actual[k], r);
        double q = orc_q2_factor(1, size, desired, rot, 0);
        if (best < q) {
            best = q;
            argmax = angle;
        }
    }
    free(rot);
    return argmax;
}

/* ScalarEqualizer::estimate: sum conj(a) d / sum |a|^2   scalar_equalizer.cpp:22-33 */
orc_c128 orc_scalar_estimate(int size, const orc_c128* desired, const orc_c128* actual) {
    double xx = 0;
    orc_c128 xy = {0, 0};
    for (int i = 0; i < size; ++i) xx += c_norm(actual[i]);
    for (int i = 0; i < size; ++i) xy = c_add(xy, c_mul(c_conj(actual[i]), desired[i]));
    orc_c128 r = {xy.re / xx, xy.im / xx};
    return r;
}

/* HomothetyEqualizer::estimate: |sum conj(a) d / sum |a|^2|   homothety_equalizer.cpp:22-33 */
double orc_homothety_estimate(int size, const orc_c128* desired, const orc_c128* actual) {
    orc_c128 r = orc_scalar_estimate(size, desired, actual);
    return hypot(r.re, r.im);
}

/* ------------------------------------------------------------------ */
/* Triplet equalizer                           sino.cpp                 */
/* ------------------------------------------------------------------ */
/* SINO::term: A_mode(k+m) * sum_p A_p(k+n) conj(A_p(k+m+n))   sino.cpp:161-171 */
void orc_sino_term(int modes, int samples, const orc_c128* f, int mode, int k, int m, int n,
                   orc_c128* out) {
    orc_c128 t = {0, 0};
    for (int p = 0; p < modes; ++p)
        t = c_add(t, c_mul(f[(long)p * samples + k + n], c_conj(f[(long)p * samples + k + m + n])));
    *out = c_mul(f[(long)mode * samples + k + m], t);
}

/* SINO::cubic: weights (2R+1)^2 row-major, W[(m+R)*(2R+1) + n+R]; edges 0
                                                   sino.cpp:203-220 */
void orc_sino_cubic(int radius, const orc_c128* weights, int modes, int samples,
                    const orc_c128* in, orc_c128* out) {
    int w = 2 * radius + 1;
    memset(out, 0, sizeof(orc_c128) * (size_t)modes * samples);
    for (int mode = 0; mode < modes; ++mode)
        for (int k = radius; k < samples - radius; ++k) {
            orc_c128 acc = {0, 0};
            for (int m = -radius; m <= radius; ++m) {
                int b = radius - abs(m);
                for (int n = -b; n <= b; ++n) {
                    orc_c128 t;
                    orc_sino_term(modes, samples, in, mode, k, m, n, &t);
                    acc = c_add(acc, c_mul(weights[(m + radius) * w + n + radius], t));
                }
            }
            out[(long)mode * samples + k] = acc;
        }
}

int orc_sino_n_features(int radius, int type) {
    int n = radius * (radius + 2) / 2 + 1;
    return type == 0 ? n + 1 : n;
}

/* SINO::train feature/response build (folded symmetric set)  sino.cpp:53-103.
   features: column-major rows x n_features, rows = modes*(samples-2R). */
void orc_sino_features(int radius, int type, orc_c128 linear_weight, int modes, int samples,
                       const orc_c128* tx, const orc_c128* rx, orc_c128* features,
                       orc_c128* responses) {
    int per_mode = samples - 2 * radius;
    long rows = (long)per_mode * modes;
    int shift = radius;
    for (int mode = 0; mode < modes; ++mode)
        for (int s = 0; s < per_mode; ++s) {
            long row = s + (long)mode * per_mode;
            orc_c128 r = rx[(long)mode * samples + s + shift];
            if (type == 1) r = c_sub(r, c_mul(tx[(long)mode * samples + s + shift], linear_weight));
            responses[row] = r;
            int index = 0;
            int mb = radius / 2;
            for (int m = -mb; m <= mb; ++m) {
                for (int n = abs(m); n <= radius - abs(m); ++n) {
                    orc_c128 a, b, c, d, v;
                    int k = s + shift;
                    if (m == 0 && n == 0) {
                        orc_sino_term(modes, samples, tx, mode, k, 0, 0, &v);
                    } else if (m == n) {
                        orc_sino_term(modes, samples, tx, mode, k, m, n, &a);
                        orc_sino_term(modes, samples, tx, mode, k, -m, -n, &b);
                        v = c_add(a, b);
                    } else if (m == -n) {
                        orc_sino_term(modes, samples, tx, mode, k, m, n, &a);
                        orc_sino_term(modes, samples, tx, mode, k, n, m, &b);
                        v = c_add(a, b);
                    } else {
                        orc_sino_term(modes, samples, tx, mode, k, m, n, &a);
                        orc_sino_term(modes, samples, tx, mode, k, n, m, &b);
                        orc_sino_term(modes, samples, tx, mode, k, -m, -n, &c);
                        orc_sino_term(modes, samples, tx, mode, k, -n, -m, &d);
                        v = c_add(c_add(a, b), c_add(c, d));
                    }
                    features[row + rows * index++] = v;
                }
            }
            if (type == 0) features[row + rows * index] = tx[(long)mode * samples + s + shift];
        }
}

/* Complex least squares min ||A x - b|| by Householder QR (the textbook
   algorithm; arma::solve's LAPACK path is not pinned bit-for-bit). */
int orc_complex_lstsq(int rows, int cols, const orc_c128* a_in, const orc_c128* b_in,
                      orc_c128* x) {
    if (rows < cols) return -1;
    orc_c128* a = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows * cols);
    orc_c128* b = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows);
    orc_c128* v = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows);
    memcpy(a, a_in, sizeof(orc_c128) * (size_t)rows * cols);
    memcpy(b, b_in, sizeof(orc_c128) * (size_t)rows);
    for (int j = 0; j < cols; ++j) {
        orc_c128* col = a + (long)j * rows;
        double nrm = 0;
        for (int i = j; i < rows; ++i) nrm += c_norm(col[i]);
        nrm = sqrt(nrm);
        if (nrm == 0) continue;
        double a0 = sqrt(c_norm(col[j]));
        orc_c128 phase = {1, 0};
        if (a0 > 0) phase = c_scale(col[j], 1.0 / a0);
        orc_c128 alpha = c_scale(phase, -nrm);
        for (int i = j; i < rows; ++i) v[i] = col[i];
        v[j] = c_sub(v[j], alpha);
        double vn = 0;
        for (int i = j; i < rows; ++i) vn += c_norm(v[i]);
        if (vn == 0) continue;
        /* apply H = I - 2 v v^H / (v^H v) to columns j.. and to b */
        for (int c = j; c < cols; ++c) {
            orc_c128* cc = a + (long)c * rows;
            orc_c128 dot = {0, 0};
            for (int i = j; i < rows; ++i) dot = c_add(dot, c_mul(c_conj(v[i]), cc[i]));
            dot = c_scale(dot, 2.0 / vn);
            for (int i = j; i < rows; ++i) cc[i] = c_sub(cc[i], c_mul(v[i], dot));
        }
        orc_c128 dot = {0, 0};
        for (int i = j; i < rows; ++i) dot = c_add(dot, c_mul(c_conj(v[i]), b[i]));
        dot = c_scale(dot, 2.0 / vn);
        for (int i = j; i < rows; ++i) b[i] = c_sub(b[i], c_mul(v[i], dot));
    }
    for (int j = cols - 1; j >= 0; --j) {
        orc_c128 s = b[j];
        for (int c = j + 1; c < cols; ++c) s = c_sub(s, c_mul(a[(long)c * rows + j], x[c]));
        x[j] = c_div(s, a[(long)j * rows + j]);
    }
    free(v);
    free(b);
    free(a);
    return 0;
}

/* z-score the feature columns (std with 1/N, arma stddev norm_type 1),
   centre the responses, least squares, de-standardise, intercept
   (sino.cpp:105-122, wdm_sino.cpp:84-100).  feat and resp are overwritten. */
static int zscore_lstsq(long rows, int nf, orc_c128* feat, orc_c128* resp, orc_c128* raw,
                        orc_c128* intercept) {
    orc_c128* mean = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)nf);
    double* sd = (double*)malloc(sizeof(double) * (size_t)nf);
    for (int j = 0; j < nf; ++j) {
        orc_c128 mu = {0, 0};
        for (long i = 0; i < rows; ++i) mu = c_add(mu, feat[i + rows * j]);
        mu = c_scale(mu, 1.0 / rows);
        double var = 0;
        for (long i = 0; i < rows; ++i) var += c_norm(c_sub(feat[i + rows * j], mu));
        mean[j] = mu;
        sd[j] = sqrt(var / rows);
    }
    orc_c128 ymean = {0, 0};
    for (long i = 0; i < rows; ++i) ymean = c_add(ymean, resp[i]);
    ymean = c_scale(ymean, 1.0 / rows);
    for (long i = 0; i < rows; ++i) resp[i] = c_sub(resp[i], ymean);
    for (int j = 0; j < nf; ++j)
        for (long i = 0; i < rows; ++i)
            feat[i + rows * j] = c_scale(c_sub(feat[i + rows * j], mean[j]), 1.0 / sd[j]);
    int rc = orc_complex_lstsq((int)rows, nf, feat, resp, raw);
    for (int j = 0; j < nf; ++j) raw[j] = c_scale(raw[j], 1.0 / sd[j]);
    orc_c128 ic = ymean;
    for (int j = 0; j < nf; ++j) ic = c_sub(ic, c_mul(raw[j], mean[j]));
    *intercept = ic;
    free(sd);
    free(mean);
    return rc;
}

/* SINO::train: z-score (std with 1/N, arma stddev norm_type 1), centre,
   least squares, de-standardise, intercept, unfold into the symmetric
   (2R+1)^2 weights                                sino.cpp:105-153 */
int orc_sino_train(int radius, int type, orc_c128 linear_weight_in, int modes, int samples,
                   const orc_c128* tx, const orc_c128* rx, orc_c128* weights,
                   orc_c128* intercept, orc_c128* linear_weight) {
    int nf = orc_sino_n_features(radius, type);
    long rows = (long)modes * (samples - 2 * radius);
    orc_c128* feat = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows * nf);
    orc_c128* resp = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows);
    orc_c128* raw = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)nf);
    orc_sino_features(radius, type, linear_weight_in, modes, samples, tx, rx, feat, resp);
    int rc = zscore_lstsq(rows, nf, feat, resp, raw, intercept);
    int w = 2 * radius + 1;
    memset(weights, 0, sizeof(orc_c128) * (size_t)w * w);
#define W_AT(i, j) weights[(i) * w + (j)]
    int index = 0;
    int mb = radius / 2;
    for (int m = -mb; m <= mb; ++m)
        for (int n = abs(m); n <= radius - abs(m); ++n) {
            int R = radius;
            if (m == 0 && n == 0) {
                W_AT(R, R) = raw[index++];
            } else if (m == n) {
                W_AT(R + m, R + n) = raw[index];
                W_AT(R - m, R - n) = raw[index++];
            } else if (m == -n) {
                W_AT(R + m, R + n) = raw[index];
                W_AT(R + n, R + m) = raw[index++];
            } else {
                W_AT(R + m, R + n) = raw[index];
                W_AT(R + n, R + m) = raw[index];
                W_AT(R - m, R - n) = raw[index];
                W_AT(R - n, R - m) = raw[index++];
            }
        }
#undef W_AT
    if (type == 0) {
        *linear_weight = raw[index];
    } else if (type == 2) {
        linear_weight->re = 0;
        linear_weight->im = 0;
    } else {
        *linear_weight = linear_weight_in;
    }
    free(raw);
    free(resp);
    free(feat);
    return rc;
}

/* ------------------------------------------------------------------ */
/* WdmPPE (6 envelope rows = ch*2 + pol, COI rows 2..3)  wdm_sino.cpp   */
/* C, D row-major: C[(c+m)*(2c+1) + (c+n)] = C_(c+m, c+n)               */
/* ------------------------------------------------------------------ */
#define A6(r, k) A6[(long)(r) * samples + (k)]
/* WdmPPE::T                                        wdm_sino.cpp:155-164 */
static orc_c128 wdm_T(const orc_c128* A6, int samples, int pol, int k, int m, int n) {
    orc_c128 t = c_add(c_mul(A6(2, k + n), c_conj(A6(2, k + m + n))),
                       c_mul(A6(3, k + n), c_conj(A6(3, k + m + n))));
    return c_mul(A6(2 + pol, k + m), t);
}
/* WdmPPE::T1                                       wdm_sino.cpp:166-182 */
static orc_c128 wdm_T1(const orc_c128* A6, int samples, int pol, int k, int m, int n) {
    orc_c128 t;
    if (pol == 0) {
        t = c_scale(c_mul(c_mul(A6(2, k + m), A6(0, k + n)), c_conj(A6(0, k + m + n))), 2.0);
        t = c_add(t, c_mul(c_mul(A6(2, k + m), A6(1, k + n)), c_conj(A6(1, k + m + n))));
        t = c_add(t, c_mul(c_mul(A6(3, k + m), A6(0, k + n)), c_conj(A6(1, k + m + n))));
    } else {
        t = c_scale(c_mul(c_mul(A6(3, k + m), A6(1, k + n)), c_conj(A6(1, k + m + n))), 2.0);
        t = c_add(t, c_mul(c_mul(A6(3, k + m), A6(0, k + n)), c_conj(A6(0, k + m + n))));
        t = c_add(t, c_mul(c_mul(A6(2, k + m), A6(1, k + n)), c_conj(A6(0, k + m + n))));
    }
    return t;
}
/* WdmPPE::T3                                       wdm_sino.cpp:184-200 */
static orc_c128 wdm_T3(const orc_c128* A6, int samples, int pol, int k, int m, int n) {
    orc_c128 t;
    if (pol == 0) {
        t = c_scale(c_mul(c_mul(A6(2, k + m), A6(4, k + n)), c_conj(A6(4, k + m + n))), 2.0);
        t = c_add(t, c_mul(c_mul(A6(2, k + m), A6(5, k + n)), c_conj(A6(5, k + m + n))));
        t = c_add(t, c_mul(c_mul(A6(3, k + m), A6(4, k + n)), c_conj(A6(5, k + m + n))));
    } else {
        t = c_scale(c_mul(c_mul(A6(3, k + m), A6(5, k + n)), c_conj(A6(5, k + m + n))), 2.0);
        t = c_add(t, c_mul(c_mul(A6(3, k + m), A6(4, k + n)), c_conj(A6(4, k + m + n))));
        t = c_add(t, c_mul(c_mul(A6(2, k + m), A6(5, k + n)), c_conj(A6(4, k + m + n))));
    }
    return t;
}
#undef A6

/* WdmPPE::spm                                      wdm_sino.cpp:256-274 */
void orc_wdmppe_spm(int c_radius, const orc_c128* C, int samples, const orc_c128* A6,
                    orc_c128* out2) {
    int w = 2 * c_radius + 1;
    memset(out2, 0, sizeof(orc_c128) * 2 * (size_t)samples);
    for (int k = c_radius; k < samples - c_radius; ++k)
        for (int pol = 0; pol < 2; ++pol) {
            orc_c128 acc = {0, 0};
            for (int n = -c_radius; n <= c_radius; ++n) {
                int b = c_radius - abs(n);
                for (int m = -b; m <= b; ++m)
                    acc = c_add(acc, c_mul(C[(c_radius + m) * w + c_radius + n],
                                           wdm_T(A6, samples, pol, k, m, n)));
            }
            out2[(long)pol * samples + k] = acc;
        }
}

/* WdmPPE::xpm                                      wdm_sino.cpp:276-295 */
void orc_wdmppe_xpm(int d_radius, const orc_c128* D, int samples, const orc_c128* A6,
                    orc_c128* out2) {
    int w = 2 * d_radius + 1;
    memset(out2, 0, sizeof(orc_c128) * 2 * (size_t)samples);
    for (int k = d_radius; k < samples - d_radius; ++k)
        for (int pol = 0; pol < 2; ++pol) {
            orc_c128 acc = {0, 0};
            for (int m = -d_radius; m <= d_radius; ++m) {
                int b = d_radius - abs(m);
                for (int n = -b; n <= b; ++n)
                    acc = c_add(acc, c_mul(D[(d_radius + m) * w + d_radius + n],
                                           c_add(wdm_T1(A6, samples, pol, k, m, n),
                                                 wdm_T3(A6, samples, pol, k, m, -n))));
            }
            out2[(long)pol * samples + k] = acc;
        }
}

int orc_wdmppe_n_features(int c_radius, int d_radius) {
    int spm = c_radius < 0 ? 0 : c_radius * (c_radius + 2) / 2 + 1;
    int xpm = d_radius < 0 ? 0 : 2 * d_radius * d_radius + 2 * d_radius + 1;
    return spm + xpm + 1;
}

/* WdmPPE::train features and responses            wdm_sino.cpp:29-81.
   Rows pol-major: row = pol * (samples - 2r) + (k - r), r = max(c, d);
   features column-major rows x n_features: the folded SPM set (as SINO,
   on the COI pair), XPM T1(m, n) + T3(m, -n) over the d-diamond (m outer),
   then the COI sample itself.  responses: rx rows 2..3. */
void orc_wdmppe_features(int c_radius, int d_radius, int samples, const orc_c128* tx6,
                         const orc_c128* rx6, orc_c128* features, orc_c128* responses) {
    int r = c_radius > d_radius ? c_radius : d_radius;
    long per_pol = samples - 2 * r;
    long rows = 2 * per_pol;
    for (int pol = 0; pol < 2; ++pol)
        for (int k = r; k < samples - r; ++k) {
            long row = pol * per_pol + (k - r);
            responses[row] = rx6[(long)(2 + pol) * samples + k];
            int index = 0;
            int mb = c_radius / 2;
            for (int m = -mb; m <= mb; ++m)
                for (int n = abs(m); n <= c_radius - abs(m); ++n) {
                    orc_c128 v;
                    if (m == 0 && n == 0) {
                        v = wdm_T(tx6, samples, pol, k, 0, 0);
                    } else if (m == n) {
                        v = c_add(wdm_T(tx6, samples, pol, k, m, n), wdm_T(tx6, samples, pol, k, -m, -n));
                    } else if (m == -n) {
                        v = c_add(wdm_T(tx6, samples, pol, k, m, n), wdm_T(tx6, samples, pol, k, n, m));
                    } else {
                        v = c_add(c_add(wdm_T(tx6, samples, pol, k, m, n), wdm_T(tx6, samples, pol, k, n, m)),
                                  c_add(wdm_T(tx6, samples, pol, k, -m, -n), wdm_T(tx6, samples, pol, k, -n, -m)));
                    }
                    features[row + rows * index++] = v;
                }
            for (int m = -d_radius; m <= d_radius; ++m) {
                int b = d_radius - abs(m);
                for (int n = -b; n <= b; ++n)
                    features[row + rows * index++] =
                        c_add(wdm_T1(tx6, samples, pol, k, m, n), wdm_T3(tx6, samples, pol, k, m, -n));
            }
            features[row + rows * index] = tx6[(long)(2 + pol) * samples + k];
        }
}

/* WdmPPE::train: z-score, least squares, de-standardise, unfold C (the SINO
   fold) and D (the diamond), linear weight last     wdm_sino.cpp:84-141.
   C (2c+1)^2 and D (2d+1)^2 row-major, written when the radius is >= 0. */
int orc_wdmppe_train(int c_radius, int d_radius, int samples, const orc_c128* tx6, const orc_c128* rx6,
                     orc_c128* C, orc_c128* D, orc_c128* intercept, orc_c128* linear_weight) {
    int r = c_radius > d_radius ? c_radius : d_radius;
    int nf = orc_wdmppe_n_features(c_radius, d_radius);
    long rows = 2L * (samples - 2 * r);
    if (rows < nf) return -1;
    orc_c128* feat = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows * nf);
    orc_c128* resp = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)rows);
    orc_c128* raw = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)nf);
    orc_wdmppe_features(c_radius, d_radius, samples, tx6, rx6, feat, resp);
    int rc = zscore_lstsq(rows, nf, feat, resp, raw, intercept);
    int index = 0;
    if (c_radius >= 0) {
        int w = 2 * c_radius + 1, R = c_radius;
        memset(C, 0, sizeof(orc_c128) * (size_t)w * w);
        int mb = c_radius / 2;
        for (int m = -mb; m <= mb; ++m)
            for (int n = abs(m); n <= c_radius - abs(m); ++n) {
                orc_c128 v = raw[index++];
                C[(R + m) * w + R + n] = v;
                if (m == 0 && n == 0) continue;
                if (m == n) {
                    C[(R - m) * w + R - n] = v;
                } else if (m == -n) {
                    C[(R + n) * w + R + m] = v;
                } else {
                    C[(R + n) * w + R + m] = v;
                    C[(R - m) * w + R - n] = v;
                    C[(R - n) * w + R - m] = v;
                }
            }
    }
    if (d_radius >= 0) {
        int w = 2 * d_radius + 1;
        memset(D, 0, sizeof(orc_c128) * (size_t)w * w);
        for (int m = -d_radius; m <= d_radius; ++m) {
            int b = d_radius - abs(m);
            for (int n = -b; n <= b; ++n) D[(d_radius + m) * w + d_radius + n] = raw[index++];
        }
    }
    *linear_weight = raw[index];
    free(raw);
    free(resp);
    free(feat);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Learned few-step DBP (no reference counterpart).                    */
/*   step j: y(t)   = sum_k h_{j,k} x((t - k + c) mod N),  c = K/2      */
/*           p(t)   = |y(t)|^2                                          */
/*           phi(t) = sum_m g_{j,m} p((t - m + P) mod N),  P = KP/2     */
/*           z(t)   = y(t) exp(i phi(t))                                */
/*   loss   = (1/(M N)) sum_{mode, t} |z_J(t) - target(t)|^2            */
/* With M = 2 modes (Manakov) the FIR acts on each mode and p(t) is the */
/* intensity summed over the modes, so the phase is common to both.     */
/* KP = 1 is the plain Kerr step exp(i theta_j |y|^2).  Gradients are   */
/* Wirtinger dL/d(conj h) for the FIR taps and dL/dg for the real       */
/* intensity taps:  q = -Im(conj(G_z) z),  dL/dphi = 2 q,                */
/*   dL/dg_m = sum_t 2 q(t) p(t - m + P),                                */
/*   dL/dp(s) = sum_m g_m 2 q(s + m - P),                                */
/*   G_y = G_z e^{-i phi} + (dL/dp) y.                                   */
/* ------------------------------------------------------------------ */
static int wrap(int i, int n) { return ((i % n) + n) % n; }

static void ldbp_fir(int taps, const orc_c128* h, int n, const orc_c128* x, orc_c128* y) {
    int c = taps / 2;
    for (int t = 0; t < n; ++t) {
        orc_c128 acc = {0, 0};
        for (int k = 0; k < taps; ++k) acc = c_add(acc, c_mul(h[k], x[wrap(t - k + c, n)]));
        y[t] = acc;
    }
}

/* phi(t) of one step from y [modes][n]: the intensity is summed over modes */
static void ldbp_phase(int ktaps, const double* g, int modes, int n, const orc_c128* y, double* phi) {
    int P = ktaps / 2;
    for (int t = 0; t < n; ++t) {
        double f = 0;
        for (int m = 0; m < ktaps; ++m) {
            double p = 0;
            for (int md = 0; md < modes; ++md) p += c_norm(y[(long)md * n + wrap(t - m + P, n)]);
            f += g[m] * p;
        }
        phi[t] = f;
    }
}

double orc_ldbp_forward(int steps, int taps, int ktaps, int modes, const orc_c128* fir, const double* kerr, int n,
                        const orc_c128* in, orc_c128* out, const orc_c128* target) {
    long mn = (long)modes * n;
    orc_c128* x = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn);
    orc_c128* y = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn);
    double* phi = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(x, in, sizeof(orc_c128) * (size_t)mn);
    for (int j = 0; j < steps; ++j) {
        for (int md = 0; md < modes; ++md) ldbp_fir(taps, fir + (long)j * taps, n, x + (long)md * n, y + (long)md * n);
        ldbp_phase(ktaps, kerr + (long)j * ktaps, modes, n, y, phi);
        for (long i = 0; i < mn; ++i) x[i] = c_mul(y[i], i_exp(phi[i % n]));
    }
    double loss = 0;
    if (target)
        for (long i = 0; i < mn; ++i) loss += c_norm(c_sub(x[i], target[i]));
    if (out) memcpy(out, x, sizeof(orc_c128) * (size_t)mn);
    free(phi);
    free(y);
    free(x);
    return loss / mn;
}

double orc_ldbp_backward(int steps, int taps, int ktaps, int modes, const orc_c128* fir, const double* kerr, int n,
                         const orc_c128* in, const orc_c128* target, orc_c128* grad_fir, double* grad_kerr) {
    /* forward with the per-step FIR inputs and outputs kept */
    long mn = (long)modes * n;
    orc_c128* xs = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn * (steps + 1));
    orc_c128* ys = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn * steps);
    double* phi = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(xs, in, sizeof(orc_c128) * (size_t)mn);
    for (int j = 0; j < steps; ++j) {
        orc_c128* x = xs + (long)j * mn;
        orc_c128* y = ys + (long)j * mn;
        for (int md = 0; md < modes; ++md) ldbp_fir(taps, fir + (long)j * taps, n, x + (long)md * n, y + (long)md * n);
        ldbp_phase(ktaps, kerr + (long)j * ktaps, modes, n, y, phi);
        for (long i = 0; i < mn; ++i) xs[(long)(j + 1) * mn + i] = c_mul(y[i], i_exp(phi[i % n]));
    }
    double loss = 0;
    orc_c128* g = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn);
    orc_c128* gy = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)mn);
    for (long i = 0; i < mn; ++i) {
        orc_c128 e = c_sub(xs[(long)steps * mn + i], target[i]);
        loss += c_norm(e);
        g[i] = c_scale(e, 1.0 / mn);
    }
    int c = taps / 2, P = ktaps / 2;
    for (int j = steps - 1; j >= 0; --j) {
        const orc_c128* x = xs + (long)j * mn;
        const orc_c128* y = ys + (long)j * mn;
        const orc_c128* z = xs + (long)(j + 1) * mn;
        const double* gk = kerr + (long)j * ktaps;
        ldbp_phase(ktaps, gk, modes, n, y, phi);
        /* q = -sum_modes Im(conj(G_z) z): the phase is common to the modes */
        for (int t = 0; t < n; ++t) {
            q[t] = 0;
            for (int md = 0; md < modes; ++md) q[t] -= c_mul(c_conj(g[(long)md * n + t]), z[(long)md * n + t]).im;
        }
        for (int m = 0; m < ktaps; ++m) {
            double acc = 0;
            for (int t = 0; t < n; ++t) {
                double p = 0;
                for (int md = 0; md < modes; ++md) p += c_norm(y[(long)md * n + wrap(t - m + P, n)]);
                acc += 2.0 * q[t] * p;
            }
            grad_kerr[(long)j * ktaps + m] = acc;
        }
        for (int s = 0; s < n; ++s) {
            double dp = 0;
            for (int m = 0; m < ktaps; ++m) dp += gk[m] * 2.0 * q[wrap(s + m - P, n)];
            for (int md = 0; md < modes; ++md) {
                long i = (long)md * n + s;
                gy[i] = c_add(c_mul(g[i], i_exp(-phi[s])), c_scale(y[i], dp));
            }
        }
        /* FIR: G_h_k = sum_modes sum_t G_y(t) conj(x(t-k+c));  G_x(s) = sum_k conj(h_k) G_y(s+k-c) */
        const orc_c128* h = fir + (long)j * taps;
        for (int k = 0; k < taps; ++k) {
            orc_c128 acc = {0, 0};
            for (int md = 0; md < modes; ++md)
                for (int t = 0; t < n; ++t)
                    acc = c_add(acc, c_mul(gy[(long)md * n + t], c_conj(x[(long)md * n + wrap(t - k + c, n)])));
            grad_fir[(long)j * taps + k] = acc;
        }
        for (int md = 0; md < modes; ++md)
            for (int s = 0; s < n; ++s) {
                orc_c128 acc = {0, 0};
                for (int k = 0; k < taps; ++k)
                    acc = c_add(acc, c_mul(c_conj(h[k]), gy[(long)md * n + wrap(s + k - c, n)]));
                g[(long)md * n + s] = acc;
            }
    }
    free(gy);
    free(g);
    free(q);
    free(phi);
    free(ys);
    free(xs);
    return loss / mn;
}

/* ---- triplet MLP post-equaliser (no reference counterpart; the triplet is
   the term of sino.cpp:154-168 with periodic indices) ---- */

enum { TM_R = 16, TM_W = 33, TM_T = 1089, TM_H = 128, TM_KF = 2240 };

int orc_tmlp_n_params(void) { return TM_H * TM_KF + TM_H + 2 * TM_H + 2; }

/* double -> float -> bfloat16 (round to nearest even), as a double */
static double tm_bf16(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

/* features of symbol s of mode md of one waveform x [modes][S] */
static void tm_features(int modes, int S, const orc_c128* x, int md, int s, int bf16, double* f) {
    for (int t = 0; t < TM_T; ++t) {
        int m = t / TM_W - TM_R, n = t % TM_W - TM_R;
        orc_c128 acc = {0, 0};
        for (int p = 0; p < modes; ++p)
            acc = c_add(acc, c_mul(x[(long)p * S + wrap(s + n, S)], c_conj(x[(long)p * S + wrap(s + m + n, S)])));
        orc_c128 T = c_mul(x[(long)md * S + wrap(s + m, S)], acc);
        f[2 * t] = bf16 ? tm_bf16(T.re) : T.re;
        f[2 * t + 1] = bf16 ? tm_bf16(T.im) : T.im;
    }
}

/* one row: pre-activations, hidden, output correction */
static void tm_row(const double* w1, const double* b1, const double* w2, const double* b2, const double* f,
                   double* pre, double* h, double* o) {
    o[0] = b2[0];
    o[1] = b2[1];
    for (int j = 0; j < TM_H; ++j) {
        double a = 0;
        const double* wj = w1 + (long)j * TM_KF;
        for (int k = 0; k < 2 * TM_T; ++k) a += wj[k] * f[k];
        pre[j] = a + b1[j];
        h[j] = pre[j] > 0 ? pre[j] : 0;
        o[0] += w2[j] * h[j];
        o[1] += w2[TM_H + j] * h[j];
    }
}

static double* tm_w1(const double* params, int bf16) {
    long n = (long)TM_H * TM_KF;
    double* w = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) w[i] = bf16 ? tm_bf16(params[i]) : params[i];
    return w;
}

double orc_tmlp_forward(int modes, int symbols, int batch, const double* params, int bf16,
                        const orc_c128* in, orc_c128* out, const orc_c128* target) {
    const double* b1 = params + (long)TM_H * TM_KF;
    const double* w2 = b1 + TM_H;
    const double* b2 = w2 + 2 * TM_H;
    double* w1 = tm_w1(params, bf16);
    double f[2 * TM_T], pre[TM_H], h[TM_H], o[2];
    double loss = 0;
    long rows = (long)batch * modes * symbols;
    for (int b = 0; b < batch; ++b) {
        const orc_c128* x = in + (long)b * modes * symbols;
        for (int md = 0; md < modes; ++md)
            for (int s = 0; s < symbols; ++s) {
                long r = ((long)b * modes + md) * symbols + s;
                tm_features(modes, symbols, x, md, s, bf16, f);
                tm_row(w1, b1, w2, b2, f, pre, h, o);
                orc_c128 y = {in[r].re + o[0], in[r].im + o[1]};
                if (out) out[r] = y;
                if (target) {
                    double e0 = y.re - target[r].re, e1 = y.im - target[r].im;
                    loss += e0 * e0 + e1 * e1;
                }
            }
    }
    free(w1);
    return loss / rows;
}

double orc_tmlp_backward(int modes, int symbols, int batch, const double* params, int bf16,
                         const orc_c128* in, const orc_c128* target, double* grad) {
    const double* b1 = params + (long)TM_H * TM_KF;
    const double* w2 = b1 + TM_H;
    const double* b2 = w2 + 2 * TM_H;
    double* gw1 = grad;
    double* gb1 = gw1 + (long)TM_H * TM_KF;
    double* gw2 = gb1 + TM_H;
    double* gb2 = gw2 + 2 * TM_H;
    memset(grad, 0, sizeof(double) * orc_tmlp_n_params());
    double* w1 = tm_w1(params, bf16);
    double f[2 * TM_T], pre[TM_H], h[TM_H], o[2];
    double loss = 0;
    long rows = (long)batch * modes * symbols;
    for (int b = 0; b < batch; ++b) {
        const orc_c128* x = in + (long)b * modes * symbols;
        for (int md = 0; md < modes; ++md)
            for (int s = 0; s < symbols; ++s) {
                long r = ((long)b * modes + md) * symbols + s;
                tm_features(modes, symbols, x, md, s, bf16, f);
                tm_row(w1, b1, w2, b2, f, pre, h, o);
                double e0 = in[r].re + o[0] - target[r].re, e1 = in[r].im + o[1] - target[r].im;
                loss += e0 * e0 + e1 * e1;
                double d0 = 2 * e0 / rows, d1 = 2 * e1 / rows;
                gb2[0] += d0;
                gb2[1] += d1;
                for (int j = 0; j < TM_H; ++j) {
                    gw2[j] += d0 * h[j];
                    gw2[TM_H + j] += d1 * h[j];
                    double dp = pre[j] > 0 ? w2[j] * d0 + w2[TM_H + j] * d1 : 0;
                    gb1[j] += dp;
                    if (bf16) dp = tm_bf16(dp);
                    double* gj = gw1 + (long)j * TM_KF;
                    for (int k = 0; k < 2 * TM_T; ++k) gj[k] += dp * f[k];
                }
            }
    }
    free(w1);
    return loss / rows;
}
