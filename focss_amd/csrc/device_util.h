// Device helpers shared by the SSFM kernels: the counter-based Gaussian
// generator used for ASE, and workgroup reductions.
#pragma once
#include "cx.h"
#include "trig.h"

namespace focss_hip {

constexpr double kPi = 3.14159265358979323846264338;  // definitions.h:14

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Box-Muller pair for element `index` of ASE stream `stream` (the span
// index).  Replaces std::normal_distribution (amplifier.cpp:25-27).
__device__ __forceinline__ void gauss_pair(unsigned long long seed, uint32_t stream,
                                           unsigned long long index, double& z0, double& z1) {
    uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0x0A5E0A5Eu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double v1 = 1.0 - u53(c[0], c[1]);
    double v2 = u53(c[2], c[3]);
    double r = sqrt(-2.0 * log(v1));
    double s, co;
    sincos(2.0 * kPi * v2, &s, &co);
    z0 = r * co;
    z1 = r * s;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of a time op's n per-tile power sums in a fixed order (lane-strided,
// then the xor butterfly, after which every lane holds the same bits): plans
// formed from it are run-to-run deterministic whatever order the tiles
// finished in.  Every lane of the wave calls it.
__device__ inline double tile_sum(const double* p, int n) {
    double s = 0;
    for (int t = threadIdx.x & 63; t < n; t += 64) s += p[t];
    return wave_sum(s);
}


// Workgroup max and sum; every thread gets the result.  red needs
// 2 * (NT / 64) doubles of shared memory.
template <int NT>
__device__ __forceinline__ void block_max_sum(double& mx, double& sm, double* red) {
    constexpr int NW = NT / 64;
    mx = wave_max(mx);
    sm = wave_sum(sm);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        red[wid] = mx;
        red[NW + wid] = sm;
    }
    __syncthreads();
    mx = red[0];
    sm = red[NW];
#pragma unroll
    for (int i = 1; i < NW; ++i) {
        mx = fmax(mx, red[i]);
        sm += red[NW + i];
    }
    __syncthreads();
}

// exp(i phase) for the dispersion multiply.  FP64 path: sincos in FP64.
// FP32 path: phase formed in FP64, reduced mod 2 pi in FP64, then rounded.
template <typename R>
__device__ __forceinline__ cx<R> phase_factor(double phase, double mag);

template <>
__device__ __forceinline__ cx<double> phase_factor<double>(double phase, double mag) {
    double s, c;
    sincos_fast(phase, &s, &c);
    return mk<double>(c * mag, s * mag);
}

template <>
__device__ __forceinline__ cx<float> phase_factor<float>(double phase, double mag) {
    const double two_pi = 6.283185307179586476925286766559;
    double red = phase - two_pi * rint(phase * (1.0 / two_pi));
    float s, c;
    sincos_fast((float)red, &s, &c);
    float m = (float)mag;
    return mk<float>(c * m, s * m);
}

// Kerr rotation exp(i coef P), computed in the field's precision.  When every
// lane of the wave has |phase| <= 2^-8 (always so on an adaptive grid, whose
// steps bound the phase by phi_max) the three-term series replaces the
// reduced sin/cos, within 1/16 the short series; the choice is uniform over
// the wave.
template <typename R>
__device__ __forceinline__ cx<R> kerr_factor(R ph) {
    R s, c;
    if (__all(fabs(ph) <= (R)0.00390625))
        sincos_tiny(ph, &s, &c);
    else if (__all(fabs(ph) <= (R)0.0625))
        sincos_small(ph, &s, &c);
    else
        sincos_fast(ph, &s, &c);
    return mk<R>(c, s);
}

}  // namespace focss_hip
