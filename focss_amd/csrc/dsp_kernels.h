// Transmitter / receiver DSP around the link program, on the device
// (src/focss/processing/modulation.cpp:86-96,248-277, src/focss/field.cpp:285-376).
// Fields are mode-major [modes][samples]; the DFTs are the library's own
// (forward scaled by 1/N, as Field::fft_inplace); these kernels are the
// pointwise and index-mapping steps between them.
#pragma once
#include "device_util.h"

namespace focss_hip {

// f[m][s] *= exp(i w s), w = 2 pi frequency dt: the phase w s is formed in
// FP64 as the reference forms it; phase_factor reduces it accurately (FP64:
// Cody-Waite modulo pi/2; FP32: modulo 2 pi in FP64 before the FP32 sin/cos)
template <typename R>
__global__ void k_frequency_shift(cx<R>* __restrict__ f, long long total, int samples, double w) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    f[i] = f[i] * phase_factor<R>(w * (double)(i % samples), 1.0);
}

// zero the spectral bins cut .. samples - cut of every mode
template <typename R>
__global__ void k_band_zero(cx<R>* __restrict__ f, long long total, int samples, int cut) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int s = (int)(i % samples);
    if (s >= cut && s <= samples - cut) f[i] = mk<R>(0, 0);
}

// spectral truncation: out[m][s] = spec[m][s] (s < cutoff) or spec[m][s + shift]
template <typename R>
__global__ void k_decimate_gather(const cx<R>* __restrict__ spec, cx<R>* __restrict__ out, int modes, int samples,
                                  int ds, int cutoff) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)modes * ds) return;
    const int m = (int)(i / ds), s = (int)(i % ds);
    const int src = s < cutoff ? s : s + (samples - ds);
    out[i] = spec[(long long)m * samples + src];
}

// f[m][s] *= h[s] for every mode
template <typename R>
__global__ void k_mul_spectrum(cx<R>* __restrict__ f, const cx<R>* __restrict__ h, long long total, int samples) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    f[i] = f[i] * h[i % samples];
}

// out (zeroed) [osf i] = in[i], flat over modes (Field::upsample on a mode-major buffer)
template <typename R>
__global__ void k_upsample(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, long long n_in, int osf) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_in) return;
    out[i * osf] = in[i];
}

// out[i] = scale in[osf i], flat over modes (Field::downsample)
template <typename R>
__global__ void k_downsample(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, long long n_out, int osf,
                             double scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    out[i] = in[i * osf] * (R)scale;
}

template <typename R>
__global__ void k_scale(cx<R>* __restrict__ f, long long total, double scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    f[i] = f[i] * (R)scale;
}

// complex128 host array converted on the device
template <typename R>
__global__ void k_from_c128(const cx<double>* __restrict__ in, cx<R>* __restrict__ out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = mk<R>((R)in[i].x, (R)in[i].y);
}

}  // namespace focss_hip
