// SSFM kernels for gfx950.
//
//  * k_resident: one workgroup owns one waveform in LDS and runs its whole
//    link program (every span, every split step) without touching HBM
//    between load and store.  Used for N*modes*sizeof(complex) <= 128 KiB.
//
//  * k_time_pass / k_spec_pass: the two-pass four-step decomposition
//    N = N1 x N2 for larger fields.  The field lives in HBM in the order
//    p = n1*N2 + n2; in the time domain that is the natural sample order.
//      time pass  (columns n2, length N1, stride N2):
//          x W_N^{-n2 k1} -> inverse length-N1 DFT -> time-domain op (Kerr,
//          ASE, span scale, power reductions) -> forward length-N1 DFT
//          -> x W_N^{n2 k1}
//      spectral pass (rows k1, length N2, contiguous):
//          forward length-N2 DFT -> x gain e^{-alpha s/2} e^{i beta2 s w^2/2} / N
//          -> inverse length-N2 DFT
//    so one split step reads and writes the field exactly twice.
#pragma once
#include "device_util.h"
#include "fft_lds.h"
#include "link.h"

namespace focss_hip {

constexpr int NT = 256;  // threads per workgroup (4 wave64)

template <typename R>
struct Tables {
    const cx<R>* bank;  // tw bank, W_{2^p}^m for p <= TW_MAX_LOG
    const cx<R>* fine;  // W_{2^20}^m, m < 1024
};

// W_N^e for N = 2^log_n <= 2^20, e in [0, N): W_{2^20}^{E}, E = e 2^{20-log_n},
// split as W_{2^10}^{E>>10} * W_{2^20}^{E&1023}.
template <typename R>
__device__ __forceinline__ cx<R> tw_big(const Tables<R>& t, unsigned e, int log_n) {
    unsigned E = (e << (20 - log_n)) & ((1u << 20) - 1);
    return t.bank[tw_offset(10) + (E >> 10)] * t.fine[E & 1023u];
}

// Logical block index grouping consecutive tiles on one XCD: the dispatcher
// deals workgroups round-robin over the 8 XCDs, so block b runs on XCD b % 8;
// tiles that share cache lines (adjacent columns) then share one L2.
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    if (nb % 8) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

struct PassArgs {
    void* field;
    const WaveStatic* statics;
    WaveState* states;   // [2][n_waves]
    Reduction* red;      // [3][n_waves]
    int* done_count;
    const int* waves;    // wave ids of this length class
    int n_waves;         // total waveforms (stride of states/red)
    int tiles_per_wave;
    int counter;         // pass counter (time pass c or spectral pass r)
    int log_n;           // log2 of samples for this class
};

// ------------------------------------------------------------------
// time-domain op on one sample (all modes)
// ------------------------------------------------------------------
template <typename R, int M>
__device__ __forceinline__ void time_op(cx<R>* a, const WaveStatic& w, const WaveState& st,
                                        long long sample, double& pmax, double& psum) {
    if (st.top == TOP_NL) {
        R p = 0;
#pragma unroll
        for (int m = 0; m < M; ++m) p += norm2(a[m]);
        cx<R> f = kerr_factor<R>((R)st.nl_coef * p);
#pragma unroll
        for (int m = 0; m < M; ++m) a[m] = a[m] * f;
    } else {
        if (st.noise_stream >= 0) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double z0, z1;
                gauss_pair(w.seed, (uint32_t)st.noise_stream,
                           (unsigned long long)(sample + (long long)m * w.samples), z0, z1);
                a[m].x += (R)(w.noise_sigma * z0);
                a[m].y += (R)(w.noise_sigma * z1);
            }
        }
        if (st.top == TOP_TRANSITION || st.top == TOP_ENTRY) {
            R sc = (R)st.t_scale;
#pragma unroll
            for (int m = 0; m < M; ++m) a[m] = a[m] * sc;
        }
    }
    double p = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) p += (double)norm2(a[m]);
    pmax = fmax(pmax, p);
    psum += p;
}

// spectral multiplier for FFT bin k of an N-point field
template <typename R>
__device__ __forceinline__ cx<R> spectral_factor(const WaveStatic& w, const SpectralOp& op,
                                                 int k, double arg, double mag) {
    int n = w.samples;
    int kk = (k <= n / 2) ? k : k - n;  // Nyquist at +N/2 (ssfm.cpp:44-47)
    double om = (double)kk * w.dw;
    double om2 = om * om;
    return phase_factor<R>(arg * om2, mag);
}

// ------------------------------------------------------------------
// resident kernel: whole link program in LDS
// ------------------------------------------------------------------
template <typename R, int N, int M>
__global__ __launch_bounds__(NT) void k_resident(PassArgs a, Tables<R> tb) {
    constexpr int LB = (ilog2<N>::value + 1) / 2;
    __shared__ cx<R> buf[padded(N * M)];
    __shared__ cx<R> tw_lo[1 << LB];
    __shared__ cx<R> tw_hi[N >> LB];
    __shared__ double redbuf[2 * (NT / 64)];
    const int w_id = a.waves[blockIdx.x];
    const WaveStatic w = a.statics[w_id];
    cx<R>* g = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    load_tw_split<R, N, LB, NT>(tw_lo, tw_hi, tb.bank);
    const TwSplit<R, LB> tw{tw_lo, tw_hi};
    for (int i = threadIdx.x; i < N * M; i += NT) {
        int m = i / N, n = i % N;
        buf[pad(n * M + m)] = g[i];
    }
    __syncthreads();
    WaveState st = initial_state(w);
    double pk = 0, sm = 0;
    // entry op
    for (int n = threadIdx.x; n < N; n += NT) {
        cx<R> v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = buf[pad(n * M + m)];
        time_op<R, M>(v, w, st, n, pk, sm);
#pragma unroll
        for (int m = 0; m < M; ++m) buf[pad(n * M + m)] = v[m];
    }
    block_max_sum<NT>(pk, sm, redbuf);
    fft_lds<R, N, M, NT, -1>(buf, tw);
    for (;;) {
        SpectralOp op = plan_spectral(w, st, pk, sm);
        if (!op.active) break;
        const double arg = 0.5 * w.beta2 * op.s;
        const double mag = op.gain * exp(-0.5 * w.alpha * op.s) / (double)N;
        for (int k = threadIdx.x; k < N; k += NT) {
            cx<R> f = spectral_factor<R>(w, op, k, arg, mag);
#pragma unroll
            for (int m = 0; m < M; ++m) buf[pad(k * M + m)] = buf[pad(k * M + m)] * f;
        }
        __syncthreads();
        fft_lds<R, N, M, NT, +1>(buf, tw);
        pk = 0;
        sm = 0;
        for (int n = threadIdx.x; n < N; n += NT) {
            cx<R> v[M];
#pragma unroll
            for (int m = 0; m < M; ++m) v[m] = buf[pad(n * M + m)];
            time_op<R, M>(v, w, st, n, pk, sm);
#pragma unroll
            for (int m = 0; m < M; ++m) buf[pad(n * M + m)] = v[m];
        }
        block_max_sum<NT>(pk, sm, redbuf);
        if (st.top == TOP_EXIT) break;
        fft_lds<R, N, M, NT, -1>(buf, tw);
    }
    for (int i = threadIdx.x; i < N * M; i += NT) {
        int m = i / N, n = i % N;
        g[i] = buf[pad(n * M + m)];
    }
    if (threadIdx.x == 0) {
        a.states[w_id] = st;
        atomicAdd(a.done_count, 1);
    }
}

// ------------------------------------------------------------------
// two-pass kernels
// ------------------------------------------------------------------
// time pass: S = C * M sequences (C columns x M modes) of length N1
template <typename R, int N1, int C, int M>
__global__ __launch_bounds__(NT) void k_time_pass(PassArgs a, Tables<R> tb) {
    constexpr int S = C * M;
    constexpr int TB = 32;  // inter-pass twiddle split k1 = hi * TB + lo
    static_assert(N1 % TB == 0, "column length must be a multiple of 32");
    __shared__ cx<R> buf[padded(N1 * S)];
    __shared__ cx<R> tw_t[N1];
    __shared__ cx<R> tw_lo[C][TB];       // W_N^{n2 lo}
    __shared__ cx<R> tw_hi[C][N1 / TB];  // W_N^{n2 TB hi}
    __shared__ double redbuf[2 * (NT / 64)];
    const int blk = xcd_block();
    const int w_id = a.waves[blk / a.tiles_per_wave];
    const int tile = blk % a.tiles_per_wave;
    const WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    if (st.top == TOP_NONE) return;
    const WaveStatic w = a.statics[w_id];
    const int N = w.samples;
    const int N2 = N / N1;
    const int n2_0 = tile * C;
    cx<R>* g = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    const bool inverse = st.top != TOP_ENTRY;
    const bool forward = st.top != TOP_EXIT;
    load_tw_full<R, N1, NT>(tw_t, tb.bank);
    const TwFull<R> tw{tw_t};
    for (int i = threadIdx.x; i < C * TB; i += NT) {
        int c = i / TB, l = i % TB;
        tw_lo[c][l] = tw_big<R>(tb, (unsigned)((n2_0 + c) * l) & (unsigned)(N - 1), a.log_n);
    }
    for (int i = threadIdx.x; i < C * (N1 / TB); i += NT) {
        int c = i / (N1 / TB), h = i % (N1 / TB);
        tw_hi[c][h] = tw_big<R>(tb, (unsigned)((n2_0 + c) * (TB * h)) & (unsigned)(N - 1), a.log_n);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N1 * S; i += NT) {
        int n1 = i / S, s = i % S;
        int m = s / C, c = s % C;
        cx<R> v = g[(long long)m * N + (long long)n1 * N2 + n2_0 + c];
        if (inverse) v = v * conj(tw_hi[c][n1 / TB] * tw_lo[c][n1 % TB]);
        buf[pad(n1 * S + s)] = v;
    }
    __syncthreads();
    if (inverse) fft_lds<R, N1, S, NT, +1>(buf, tw);
    double pk = 0, sm = 0;
    for (int i = threadIdx.x; i < N1 * C; i += NT) {
        int n1 = i / C, c = i % C;
        cx<R> v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = buf[pad(n1 * S + m * C + c)];
        time_op<R, M>(v, w, st, (long long)n1 * N2 + n2_0 + c, pk, sm);
#pragma unroll
        for (int m = 0; m < M; ++m) buf[pad(n1 * S + m * C + c)] = v[m];
    }
    block_max_sum<NT>(pk, sm, redbuf);
    if (threadIdx.x == 0) {
        Reduction* r = a.red + (a.counter % 3) * a.n_waves + w_id;
        atomicMax(&r->peak_bits, double_to_bits(pk));
        atomicAdd(&r->sum, sm);
    }
    if (forward) fft_lds<R, N1, S, NT, -1>(buf, tw);
    for (int i = threadIdx.x; i < N1 * S; i += NT) {
        int n1 = i / S, s = i % S;
        int m = s / C, c = s % C;
        cx<R> v = buf[pad(n1 * S + s)];
        if (forward) v = v * (tw_hi[c][n1 / TB] * tw_lo[c][n1 % TB]);
        g[(long long)m * N + (long long)n1 * N2 + n2_0 + c] = v;
    }
}

// spectral pass: S rows of length N2; row q = m * N1 + k1 is contiguous
template <typename R, int N2, int S>
__global__ __launch_bounds__(NT) void k_spec_pass(PassArgs a, Tables<R> tb) {
    __shared__ cx<R> buf[padded(N2 * S)];
    __shared__ cx<R> tw_t[N2];
    const int blk = xcd_block();
    const int w_id = a.waves[blk / a.tiles_per_wave];
    const int tile = blk % a.tiles_per_wave;
    WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    const int was_done = st.phase == PH_DONE;
    const WaveStatic w = a.statics[w_id];
    const Reduction rd = a.red[(a.counter % 3) * a.n_waves + w_id];
    SpectralOp op = plan_spectral(w, st, bits_to_double(rd.peak_bits), rd.sum);
    if (tile == 0 && threadIdx.x == 0) {
        a.states[((a.counter + 1) & 1) * a.n_waves + w_id] = st;
        Reduction z;
        z.peak_bits = 0;
        z.sum = 0;
        a.red[((a.counter + 1) % 3) * a.n_waves + w_id] = z;
        if (!was_done && st.phase == PH_DONE) atomicAdd(a.done_count, 1);
    }
    if (!op.active) return;
    const int N = w.samples;
    const int N1 = N / N2;
    cx<R>* g = reinterpret_cast<cx<R>*>(a.field) + w.offset + (long long)tile * S * N2;
    load_tw_full<R, N2, NT>(tw_t, tb.bank);
    const TwFull<R> tw{tw_t};
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        int s = i / N2, n2 = i % N2;
        buf[pad(n2 * S + s)] = g[i];
    }
    __syncthreads();
    fft_lds<R, N2, S, NT, -1>(buf, tw);
    const double arg = 0.5 * w.beta2 * op.s;
    const double mag = op.gain * exp(-0.5 * w.alpha * op.s) / (double)N;
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        int k2 = i / S, s = i % S;
        int q = tile * S + s;
        int k1 = q % N1;
        buf[pad(i)] = buf[pad(i)] * spectral_factor<R>(w, op, k1 + N1 * k2, arg, mag);
    }
    __syncthreads();
    fft_lds<R, N2, S, NT, +1>(buf, tw);
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        int s = i / N2, n2 = i % N2;
        g[i] = buf[pad(n2 * S + s)];
    }
}

}  // namespace focss_hip
