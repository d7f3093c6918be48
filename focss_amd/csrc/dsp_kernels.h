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

// every element times the complex scalar (re, im) (Field::operator*=,
// src/focss/field.cpp:381-392), grid-stride
template <typename R>
__global__ void k_mul_scalar(cx<R>* __restrict__ f, long long total, double re, double im) {
    const cx<R> m = mk<R>((R)re, (R)im);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x)
        f[i] = f[i] * m;
}

// complex128 host array converted on the device
template <typename R>
__global__ void k_from_c128(const cx<double>* __restrict__ in, cx<R>* __restrict__ out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = mk<R>((R)in[i].x, (R)in[i].y);
}

}  // namespace focss_hip

namespace focss_hip {

// Records of save_field / load_field (src/tasks/wdm_run.cpp:55-77,
// src/tasks/wdm_scoring.cpp:79-103) hold complex<double> in sample-major
// order [sample][envelope], envelope = channel * modes + mode; the field
// layout is envelope-major [envelope][sample].  A workgroup moves a tile of
// (IO_TILE >> te_log2) samples x (1 << te_log2) envelopes through LDS, so the
// record side is read or written in runs of whole sample rows and the field
// side in runs along samples.  Rows of the LDS tile are padded by one value.
// TO_FIELD: records -> field of precision R; otherwise field -> records.
constexpr int IO_LOG_TILE = 11;
constexpr int IO_TILE = 1 << IO_LOG_TILE;
constexpr int IO_NT = 256;

template <typename R, bool TO_FIELD>
__global__ __launch_bounds__(IO_NT) void k_envelope_transpose(const void* __restrict__ src, void* __restrict__ dst,
                                                               int samples, int envelopes, int te_log2) {
    __shared__ cx<double> tile[IO_TILE + 32];
    const int te = 1 << te_log2, ts_log2 = IO_LOG_TILE - te_log2, ts = 1 << ts_log2, row = ts + 1;
    const long long s0 = (long long)blockIdx.x * ts;
    const int e0 = blockIdx.y * te;
    if (TO_FIELD) {
        const cx<double>* rec = (const cx<double>*)src;
        for (int k = threadIdx.x; k < IO_TILE; k += IO_NT) {
            const int s = k >> te_log2, e = k & (te - 1);
            if (s0 + s < samples && e0 + e < envelopes) tile[e * row + s] = rec[(s0 + s) * envelopes + e0 + e];
        }
        __syncthreads();
        cx<R>* f = (cx<R>*)dst;
        for (int k = threadIdx.x; k < IO_TILE; k += IO_NT) {
            const int e = k >> ts_log2, s = k & (ts - 1);
            if (s0 + s < samples && e0 + e < envelopes) {
                const cx<double> v = tile[e * row + s];
                f[(long long)(e0 + e) * samples + s0 + s] = mk<R>((R)v.x, (R)v.y);
            }
        }
    } else {
        const cx<R>* f = (const cx<R>*)src;
        for (int k = threadIdx.x; k < IO_TILE; k += IO_NT) {
            const int e = k >> ts_log2, s = k & (ts - 1);
            if (s0 + s < samples && e0 + e < envelopes) {
                const cx<R> v = f[(long long)(e0 + e) * samples + s0 + s];
                tile[e * row + s] = mk<double>((double)v.x, (double)v.y);
            }
        }
        __syncthreads();
        cx<double>* rec = (cx<double>*)dst;
        for (int k = threadIdx.x; k < IO_TILE; k += IO_NT) {
            const int s = k >> te_log2, e = k & (te - 1);
            if (s0 + s < samples && e0 + e < envelopes) rec[(s0 + s) * envelopes + e0 + e] = tile[e * row + s];
        }
    }
}

// Sums behind the constant equalisers of src/focss/equalizer (phase search
// phase_shift_equalizer.cpp:30-46, complex scalar scalar_equalizer.cpp:22-33,
// real scalar homothety_equalizer.cpp:22-33), over n complex elements in FP64:
//     out[0] += sum |d|^2,  out[1] += sum |a|^2,  out[2] + i out[3] += sum conj(a) d.
constexpr int LS_NT = 256;
template <typename R>
__global__ __launch_bounds__(LS_NT) void k_linear_stats(const cx<R>* __restrict__ d, const cx<R>* __restrict__ a,
                                                        long long n, double* __restrict__ out) {
    __shared__ double red[4][LS_NT / 64];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long i = blockIdx.x * (long long)LS_NT + threadIdx.x; i < n; i += (long long)gridDim.x * LS_NT) {
        const double dr = d[i].x, di = d[i].y, ar = a[i].x, ai = a[i].y;
        s[0] += dr * dr + di * di;
        s[1] += ar * ar + ai * ai;
        s[2] += ar * dr + ai * di;
        s[3] += ar * di - ai * dr;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = s[k];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int w = 0; w < LS_NT / 64; ++w) t += red[threadIdx.x][w];
        atomicAdd(out + threadIdx.x, t);
    }
}

}  // namespace focss_hip
