// Checks focss_hip::sincos_fast against long-double libm over the ranges the
// split-step kernels use; prints the worst absolute errors.
#include <cmath>
#include <cstdio>
#include <random>
#include "trig.h"

int main() {
    std::mt19937_64 rng(12345);
    double e64 = 0, e32 = 0, e64_big = 0;
    auto check64 = [&](double x, double& e) {
        double s, c;
        focss_hip::sincos_fast(x, &s, &c);
        long double ls = sinl((long double)x), lc = cosl((long double)x);
        e = fmax(e, (double)fabsl(s - ls));
        e = fmax(e, (double)fabsl(c - lc));
    };
    auto check32 = [&](float x) {
        float s, c;
        focss_hip::sincos_fast(x, &s, &c);
        long double ls = sinl((long double)x), lc = cosl((long double)x);
        e32 = fmax(e32, (double)fabsl(s - ls));
        e32 = fmax(e32, (double)fabsl(c - lc));
    };
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int i = 0; i < 2000000; ++i) {
        double scale = std::pow(10.0, 6.0 * (u(rng) + 1.0) / 2.0 - 4.0);  // 1e-4 .. 1e2
        check64(u(rng) * scale, e64);
        check32((float)(u(rng) * scale));
    }
    for (int i = 0; i < 1000000; ++i) check64(u(rng) * 1.0e6, e64_big);
    for (int k = -2000; k <= 2000; ++k) {  // neighbourhoods of multiples of pi/2
        double x = k * 1.5707963267948966;
        for (int d = -3; d <= 3; ++d) {
            check64(std::nextafter(x, x + d), e64);
            check32((float)x + d * 1e-7f);
        }
    }
    double e_small = 0, e_small32 = 0;
    for (int i = -200000; i <= 200000; ++i) {  // the short series on |x| <= 1/16
        double x = 0.0625 * i / 200000.0;
        double ss, cc;
        focss_hip::sincos_small(x, &ss, &cc);
        e_small = fmax(e_small, fmax(fabs(ss - (double)sinl(x)), fabs(cc - (double)cosl(x))));
        float sf, cf, xf = (float)x;
        focss_hip::sincos_small(xf, &sf, &cf);
        e_small32 = fmax(e_small32, fmax(fabs(sf - (double)sinl(xf)), fabs(cf - (double)cosl(xf))));
    }
    double e_tiny = 0, e_tiny32 = 0;
    for (int i = -200000; i <= 200000; ++i) {  // the three-term series on |x| <= 2^-8
        double x = 0.00390625 * i / 200000.0;
        double ss, cc;
        focss_hip::sincos_tiny(x, &ss, &cc);
        e_tiny = fmax(e_tiny, fmax(fabs(ss - (double)sinl(x)), fabs(cc - (double)cosl(x))));
        float sf, cf, xf = (float)x;
        focss_hip::sincos_tiny(xf, &sf, &cf);
        e_tiny32 = fmax(e_tiny32, fmax(fabs(sf - (double)sinl(xf)), fabs(cf - (double)cosl(xf))));
    }
    double s, c;
    focss_hip::sincos_fast(3.0e7, &s, &c);  // library fallback
    double e_fb = fmax(fabs(s - sin(3.0e7)), fabs(c - cos(3.0e7)));
    printf("%.3e %.3e %.3e %.3e %.3e %.3e %.3e %.3e\n", e64, e64_big, e32, e_fb, e_small, e_small32, e_tiny,
           e_tiny32);
    return 0;
}
