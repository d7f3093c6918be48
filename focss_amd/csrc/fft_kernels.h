// Standalone batched DFT kernels behind Field::fft_inplace / ifft_inplace
// (src/focss/field.cpp:343-352): forward e^{-2 pi i jk/N} followed by x1/N,
// inverse unscaled, `modes` rows of N samples each.
//   N <= 8192: one workgroup per row, whole row in LDS.
//   larger N : column pass (length N1 DFT of stride-N2 columns, x W_N^{+-n2 k1},
//              from the field into a work buffer at the same positions) then
//              row pass (length N2 DFT of each contiguous row of the work
//              buffer, written back to the field in natural order k1 + N1 k2):
//              one read and one write of the field per pass.
#pragma once
#include "device_util.h"
#include "fft_lds.h"
#include "ssfm_kernels.h"

namespace focss_hip {

template <typename R, int N, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_whole(cx<R>* data, double scale, Tables<R> tb) {
    constexpr int LB = (ilog2<N>::value + 1) / 2;
    __shared__ cx<R> buf[padded(N)];
    __shared__ cx<R> tw_lo[1 << LB];
    __shared__ cx<R> tw_hi[N >> LB];
    cx<R>* g = data + (long long)blockIdx.x * N;
    load_tw_split<R, N, LB, NT>(tw_lo, tw_hi, tb.bank);
    const TwSplit<R, LB> tw{tw_lo, tw_hi};
    for (int i = threadIdx.x; i < N; i += NT) buf[pad(i)] = g[i];
    __syncthreads();
    fft_lds<R, N, 1, NT, SIGN>(buf, tw);
    const R sc = (R)scale;
    for (int i = threadIdx.x; i < N; i += NT) g[i] = buf[pad(i)] * sc;
}

template <typename R, int N1, int C, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_cols(const cx<R>* data, cx<R>* work, int log_n, int tiles_per_row,
                                                 Tables<R> tb) {
    __shared__ cx<R> buf[padded(N1 * C)];
    __shared__ cx<R> tw_t[N1];
    const int N = 1 << log_n;
    const int N2 = N / N1;
    const int row = blockIdx.x / tiles_per_row;
    const int n2_0 = (blockIdx.x % tiles_per_row) * C;
    const cx<R>* g = data + (long long)row * N;
    cx<R>* o = work + (long long)row * N;
    load_tw_full<R, N1, NT>(tw_t, tb.bank);
    const TwFull<R> tw{tw_t};
    for (int i = threadIdx.x; i < N1 * C; i += NT) {
        int n1 = i / C, c = i % C;
        buf[pad(i)] = g[(long long)n1 * N2 + n2_0 + c];
    }
    __syncthreads();
    fft_lds<R, N1, C, NT, SIGN>(buf, tw);
    for (int i = threadIdx.x; i < N1 * C; i += NT) {
        int k1 = i / C, c = i % C;
        cx<R> w = tw_big<R>(tb, (unsigned)((n2_0 + c) * k1) & (unsigned)(N - 1), log_n);
        o[(long long)k1 * N2 + n2_0 + c] = buf[pad(i)] * (SIGN < 0 ? w : conj(w));
    }
}

template <typename R, int N2, int S, int SIGN>
__global__ __launch_bounds__(NT) void k_fft_rows(const cx<R>* src, cx<R>* dst, int log_n, int tiles_per_row,
                                                 double scale, Tables<R> tb) {
    __shared__ cx<R> buf[padded(N2 * S)];
    __shared__ cx<R> tw_t[N2];
    const int N = 1 << log_n;
    const int N1 = N / N2;
    const int row = blockIdx.x / tiles_per_row;
    const int k1_0 = (blockIdx.x % tiles_per_row) * S;
    const cx<R>* g = src + (long long)row * N + (long long)k1_0 * N2;
    load_tw_full<R, N2, NT>(tw_t, tb.bank);
    const TwFull<R> tw{tw_t};
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        int s = i / N2, n2 = i % N2;
        buf[pad(n2 * S + s)] = g[i];
    }
    __syncthreads();
    fft_lds<R, N2, S, NT, SIGN>(buf, tw);
    const R sc = (R)scale;
    cx<R>* o = dst + (long long)row * N;
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        int k2 = i / S, s = i % S;
        o[(long long)(k1_0 + s) + (long long)N1 * k2] = buf[pad(i)] * sc;
    }
}

// Amplifier::add_noise (src/focss/module/amplifier.cpp:20-28): element i of
// the mode-major field gets sigma (z0 + i z1) from Philox stream `stream`.
template <typename R>
__global__ void k_add_noise(cx<R>* data, long long n, double sigma, unsigned long long seed, uint32_t stream) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double z0, z1;
        gauss_pair(seed, stream, (unsigned long long)i, z0, z1);
        cx<R> v = data[i];
        v.x += (R)(sigma * z0);
        v.y += (R)(sigma * z1);
        data[i] = v;
    }
}

}  // namespace focss_hip
