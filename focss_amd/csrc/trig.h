// sin and cos for the phase rotations of the split-step kernels.
//
// Argument reduction modulo pi/2 after Cody and Waite ("Software Manual for
// the Elementary Functions", 1980): k = rint(2x/pi), r = x - k*hi - k*lo with
// pi/2 = hi + lo.  With a fused multiply-add, x - k*hi is exact for any
// |x| <= 2^20 (both terms are multiples of ulp(hi) and |r| < 1, or Sterbenz
// applies), and the error of the lo term is below 1e-26 there.  sin and cos of
// r in [-pi/4, pi/4] come from their Taylor series, truncated where the next
// term is below 1e-19 (double) or 1e-10 (float) relative, in Horner form.
// Quadrant k mod 4 selects and negates.  Arguments beyond the reduction's
// range fall back to the device library.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define FOCSS_TRIG_HD __host__ __device__
#else
#define FOCSS_TRIG_HD
#endif

namespace focss_hip {

FOCSS_TRIG_HD inline void quadrant(int q, double sr, double cr, double* s, double* c) {
    double a = (q & 1) ? cr : sr;
    double b = (q & 1) ? sr : cr;
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

FOCSS_TRIG_HD inline void quadrant(int q, float sr, float cr, float* s, float* c) {
    float a = (q & 1) ? cr : sr;
    float b = (q & 1) ? sr : cr;
    *s = (q & 2) ? -a : a;
    *c = ((q + 1) & 2) ? -b : b;
}

// |x| <= 2^20: |error| <= ~2 ulp of 1
FOCSS_TRIG_HD inline void sincos_reduced(double x, double* s, double* c) {
    const double two_over_pi = 0.6366197723675814;
    const double pio2_hi = 1.5707963267948966;     // pi/2 rounded to double
    const double pio2_lo = 6.123233995736766e-17;  // pi/2 - pio2_hi
    const double k = rint(x * two_over_pi);
    double r = fma(-k, pio2_hi, x);
    r = fma(-k, pio2_lo, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2/5! - ... + r^14/17!)
    double ps = 2.8114572543455206e-15;
    ps = fma(ps, r2, -7.647163731819816e-13);
    ps = fma(ps, r2, 1.6059043836821613e-10);
    ps = fma(ps, r2, -2.505210838544172e-08);
    ps = fma(ps, r2, 2.7557319223985893e-06);
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    // cos r = 1 - r^2/2 + r^4 (1/4! - r^2/6! + ... + r^12/16!)
    double pc = 4.779477332387385e-14;
    pc = fma(pc, r2, -1.1470745597729725e-11);
    pc = fma(pc, r2, 2.08767569878681e-09);
    pc = fma(pc, r2, -2.755731922398589e-07);
    pc = fma(pc, r2, 2.48015873015873e-05);
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    const double cr = fma(r2 * r2, pc, fma(-0.5, r2, 1.0));
    quadrant((int)k, sr, cr, s, c);
}

// |x| <= 2^14: |error| <= ~2 ulp of 1
FOCSS_TRIG_HD inline void sincos_reduced(float x, float* s, float* c) {
    const float two_over_pi = 0.6366197466850281f;
    const float pio2_hi = 1.5707963705062866f;     // pi/2 rounded to float
    const float pio2_lo = -4.371138828673793e-08f;  // pi/2 - pio2_hi
    const float k = rintf(x * two_over_pi);
    float r = fmaf(-k, pio2_hi, x);
    r = fmaf(-k, pio2_lo, r);
    const float r2 = r * r;
    float ps = 2.7557319223985893e-06f;
    ps = fmaf(ps, r2, -0.0001984126984126984f);
    ps = fmaf(ps, r2, 0.008333333333333333f);
    ps = fmaf(ps, r2, -0.16666666666666666f);
    const float sr = fmaf(r * r2, ps, r);
    float pc = -2.755731922398589e-07f;
    pc = fmaf(pc, r2, 2.48015873015873e-05f);
    pc = fmaf(pc, r2, -0.001388888888888889f);
    pc = fmaf(pc, r2, 0.041666666666666664f);
    const float cr = fmaf(r2 * r2, pc, fmaf(-0.5f, r2, 1.0f));
    quadrant((int)k, sr, cr, s, c);
}

// |x| <= 1/16 (the per-step nonlinear phases of an adaptive grid): the series
// without reduction, truncated after x^9 (sin) and x^10 (cos); the first
// omitted terms are below 2e-21 and 8e-24
FOCSS_TRIG_HD inline void sincos_small(double x, double* s, double* c) {
    const double x2 = x * x;
    double ps = 2.7557319223985893e-06;
    ps = fma(ps, x2, -0.0001984126984126984);
    ps = fma(ps, x2, 0.008333333333333333);
    ps = fma(ps, x2, -0.16666666666666666);
    *s = fma(x * x2, ps, x);
    double pc = -2.755731922398589e-07;
    pc = fma(pc, x2, 2.48015873015873e-05);
    pc = fma(pc, x2, -0.001388888888888889);
    pc = fma(pc, x2, 0.041666666666666664);
    *c = fma(x2 * x2, pc, fma(-0.5, x2, 1.0));
}

// |x| <= 1/16: truncated after x^5 and x^6 (omitted terms below 1e-12)
FOCSS_TRIG_HD inline void sincos_small(float x, float* s, float* c) {
    const float x2 = x * x;
    *s = fmaf(x * x2, fmaf(x2, 0.008333333333333333f, -0.16666666666666666f), x);
    *c = fmaf(x2 * x2, fmaf(x2, -0.001388888888888889f, 0.041666666666666664f), fmaf(-0.5f, x2, 1.0f));
}

// |x| <= 2^-8 (every step of an adaptive grid: its phase is at most phi_max):
// sin to x^5, cos to x^4; the first omitted terms are below 3e-21 and 5e-18
FOCSS_TRIG_HD inline void sincos_tiny(double x, double* s, double* c) {
    const double x2 = x * x;
    *s = fma(x * x2, fma(x2, 0.008333333333333333, -0.16666666666666666), x);
    *c = fma(x2, fma(x2, 0.041666666666666664, -0.5), 1.0);
}

// |x| <= 2^-8: sin to x^3, cos to x^2; omitted terms below 8e-15 and 1e-11
FOCSS_TRIG_HD inline void sincos_tiny(float x, float* s, float* c) {
    const float x2 = x * x;
    *s = fmaf(x * x2, -0.16666666666666666f, x);
    *c = fmaf(-0.5f, x2, 1.0f);
}

FOCSS_TRIG_HD inline void sincos_fast(double x, double* s, double* c) {
    if (fabs(x) <= 1048576.0)
        sincos_reduced(x, s, c);
    else
        sincos(x, s, c);
}

FOCSS_TRIG_HD inline void sincos_fast(float x, float* s, float* c) {
    if (fabsf(x) <= 16384.0f)
        sincos_reduced(x, s, c);
    else
        sincosf(x, s, c);
}

}  // namespace focss_hip
