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
#include <type_traits>
#include "device_util.h"
#include "fft_lds.h"
#include "link.h"

namespace focss_hip {

constexpr int NT = 256;  // threads per workgroup (4 wave64)

// Workgroup size of a two-pass kernel with `items` 8-point work items: one
// item per thread, 64 to 512 threads.
__host__ __device__ constexpr int pass_threads(int items) {
    return items < 64 ? 64 : (items > 512 ? 512 : items);
}

template <typename R>
struct Tables {
    const cx<R>* bank;  // tw bank, W_{2^p}^m for p <= TW_MAX_LOG
    const cx<R>* fine;  // W_{2^20}^m, m < 1024
    const cx<double>* bank64;  // the same two tables in FP64
    const cx<double>* fine64;
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
    int layout_g;        // 2: row-pair interleaved intermediate layout, 1: natural
    double* partials;    // [3][n_waves][partial_stride] per-tile power sums (LOGARITHMIC grid)
    int partial_stride;
    int time_tiles;      // time-pass tiles per waveform of this class (spectral pass)
};

// Intermediate layout of the two-pass programs.  The field is an N1 x N2
// matrix (row k1 or n1, column n2) per mode; natural order puts (row, col)
// at row N2 + col.  The row-pair interleaved order (layout_g = 2) stores
// each pair of rows as one block of 2 N2 values, in chunks of W columns
// (64 bytes) alternating between the two rows:
//     (row, col) at (row / 2) 2 N2 + (col / W) 2 W + (row % 2) W + col % W.
// A time-pass tile of C >= W columns then reads and writes 2 W-value
// (128-byte) runs instead of W-value runs -- measured 4.8 - 5.2 against
// 4.1 - 4.3 TB/s for the copy alone -- while a spectral-pass tile of an even
// number of rows stays one contiguous block in either order.  So the first
// spectral pass turns natural order into interleaved order in place, every
// pass after it keeps it, and k_deinterleave restores natural order once
// the programs end.
template <typename R>
struct Interleave {
    static constexpr int W = 64 / (int)sizeof(cx<R>);
    // offset of (row, col); the caller adds the mode's base
    static __device__ __forceinline__ long long at(int row, int col, int N2) {
        return (long long)(row >> 1) * (2 * N2) + (col / W) * (2 * W) + (row & 1) * W + (col % W);
    }
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
        if (threadIdx.x == 0) record_step(w, st);
        if (!op.active) break;
        const double arg = 0.5 * w.beta2 * op.s;
        const double mag = op.gain * op.loss / (double)N;
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
// Both passes run their two DFTs as split-stage schedules (fft_lds.h): the
// first stage of the first DFT reads HBM, the last stage of the first DFT,
// the pointwise operator and the first stage of the second DFT share
// registers, and the last stage of the second DFT writes HBM, so each
// element crosses LDS 8 times per pass.  A work item is (sequence group c,
// j < L/8); it owns elements j + r L/8, r < 8.

// Inter-pass twiddles W_N^{n2 (j + r Q)}, r < 8, formed in FP64 from two bank
// lookups and products of depth <= 4, then rounded once to R: t0 = W^{n2 j}
// and q1 = W^{n2 Q} from the tables (inter_bases), then t0 q1^r (inter_form).
template <typename R>
__device__ __forceinline__ void inter_bases(const Tables<R>& tb, unsigned n2, unsigned j, unsigned Q, int log_n,
                                            cx<double>& t0, cx<double>& q1) {
    const unsigned mask = (1u << log_n) - 1u;
    const unsigned sh = 20 - log_n;
    auto big = [&](unsigned e) {
        unsigned E = (e << sh) & ((1u << 20) - 1u);
        return tb.bank64[tw_offset(10) + (E >> 10)] * tb.fine64[E & 1023u];
    };
    t0 = big((n2 * j) & mask);
    q1 = big((n2 * Q) & mask);
}

// the four table values of inter_bases, not yet multiplied
template <typename R>
__device__ __forceinline__ void inter_loads(const Tables<R>& tb, unsigned n2, unsigned j, unsigned Q, int log_n,
                                            cx<double>* t0, cx<double>* q1) {
    const unsigned mask = (1u << log_n) - 1u;
    const unsigned sh = 20 - log_n;
    const unsigned E0 = ((((n2 * j) & mask) << sh) & ((1u << 20) - 1u));
    const unsigned E1 = ((((n2 * Q) & mask) << sh) & ((1u << 20) - 1u));
    t0[0] = tb.bank64[tw_offset(10) + (E0 >> 10)];
    t0[1] = tb.fine64[E0 & 1023u];
    q1[0] = tb.bank64[tw_offset(10) + (E1 >> 10)];
    q1[1] = tb.fine64[E1 & 1023u];
}

template <typename R>
__device__ __forceinline__ void inter_form(cx<double> t0, cx<double> q1, cx<R>* t) {
    cx<double> q2 = q1 * q1;
    cx<double> q4 = q2 * q2;
    cx<double> d[8];
    d[0] = t0;
    d[1] = t0 * q1;
    d[2] = t0 * q2;
    d[3] = d[1] * q2;
    d[4] = t0 * q4;
    d[5] = d[1] * q4;
    d[6] = d[2] * q4;
    d[7] = d[3] * q4;
#pragma unroll
    for (int r = 0; r < 8; ++r) t[r] = mk<R>((R)d[r].x, (R)d[r].y);
}

template <typename R>
__device__ __forceinline__ void inter_twiddles(const Tables<R>& tb, unsigned n2, unsigned j, unsigned Q, int log_n,
                                               cx<R>* t) {
    cx<double> t0, q1;
    inter_bases<R>(tb, n2, j, Q, log_n, t0, q1);
    inter_form<R>(t0, q1, t);
}

// time pass: C columns x M modes, each column of length N1 with stride N2.
// A workgroup runs TPB consecutive tiles of one waveform; with TPB = 2 the
// second tile's loads are issued before the first tile's DFTs (barriers
// wait on LDS only, so they stay in flight).  The inter-pass twiddles are
// formed at the load and again at the store rather than held in registers
// across both DFTs where that pays (124 instead of 150 VGPRs for 512 x 4
// complex128: four waves per SIMD instead of three).
template <typename R, int N1, int C, int M, int TPB>
__global__ __launch_bounds__(pass_threads(C * N1 / 8)) void k_time_pass(PassArgs a, Tables<R> tb) {
    constexpr int S = C * M;
    constexpr int Q = N1 / 8;
    constexpr int ITEMS = C * Q;
    constexpr int TN = pass_threads(ITEMS);
    constexpr int PER = (ITEMS + TN - 1) / TN;
    static_assert(N1 >= 64, "split-stage DFT needs length >= 64");
    using TW = PassTw<R, N1>;
    __shared__ cx<R> buf[padded(N1 * S)];
    __shared__ cx<R> tw_a[TW::A];
    __shared__ cx<R> tw_b[TW::B];
    __shared__ double redbuf[2 * (TN / 64)];
    const int blk = xcd_block();
    const int groups = a.tiles_per_wave / TPB;
    const int w_id = a.waves[blk / groups];
    const int tile0 = (blk % groups) * TPB;
    const WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    if (st.top == TOP_NONE) return;
    const WaveStatic w = a.statics[w_id];
    const int N = w.samples;
    const int N2 = N / N1;
    cx<R>* g = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    const bool inverse = st.top != TOP_ENTRY;
    const bool forward = st.top != TOP_EXIT;
    // the entry pass sees the caller's natural order; rows j + r Q (Q even)
    // are r Q N2 apart in either order
    const bool il = a.layout_g == 2 && st.top != TOP_ENTRY;
    const long long rstride = (long long)Q * N2;
    auto at = [&](int j, int n2) -> long long {
        return il ? Interleave<R>::at(j, n2, N2) : (long long)j * N2 + n2;
    };
    auto load = [&](int tile, cx<R> (&dst)[PER][M][8]) {
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const long long base = at(id / C, tile * C + id % C);
#pragma unroll
                for (int m = 0; m < M; ++m) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) dst[b][m][r] = g[(long long)m * N + base + r * rstride];
                }
            }
        }
    };
    const auto tw_t = TW::template load<TN>(tw_a, tw_b, tb.bank);
    static_assert(TPB == 1 || TPB == 2, "one or two tiles per workgroup");
    // the inter-pass twiddles formed again at the store (single-mode FP64
    // columns, or two tiles) or held from the load (two modes: 88 against
    // 111 us for 1024 x 2 x 2 complex128; FP32, where eight twiddles are 16
    // registers: 78-79 against 83 us for 512 x 4 complex64)
    constexpr bool kRecompute = (M == 1 && sizeof(R) == 8) || TPB > 1;
    // (a loop here allocates fewer registers than the compile-time tile
    // calls of the spectral pass)
    cx<R> v[PER][M][8];
#pragma unroll
    for (int t = 0; t < TPB; ++t) {
        const int tile = tile0 + t;
        const int n2_0 = tile * C;
        cx<R> itw[PER][8];
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
                inter_twiddles<R>(tb, (unsigned)(n2_0 + c), (unsigned)j, (unsigned)Q, a.log_n, itw[b]);
            }
        }
        if (t == 0) load(tile, v);  // later tiles: loaded during the previous one
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
                if (inverse) {
#pragma unroll
                    for (int m = 0; m < M; ++m) {
#pragma unroll
                        for (int r = 0; r < 8; ++r) v[b][m][r] = v[b][m][r] * conj(itw[b][r]);
                        first_stage_store<R, N1, S, +1>(buf, v[b][m], m * C + c, j);
                    }
                }
            }
        }
        cx<R> vn[PER][M][8];
        if (t + 1 < TPB) load(tile + 1, vn);
        if (inverse) {
            __syncthreads();
            middle_stages<R, N1, S, TN, +1, false>(buf, tw_t);
#pragma unroll
            for (int b = 0; b < PER; ++b) {
                const int id = threadIdx.x + b * TN;
                if (ITEMS % TN == 0 || id < ITEMS) {
                    const int c = id % C, j = id / C;
#pragma unroll
                    for (int m = 0; m < M; ++m) last_stage_load<R, N1, S, +1>(buf, tw_t, v[b][m], m * C + c, j);
                }
            }
        }
        double pk = 0, sm = 0;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    cx<R> u[M];
#pragma unroll
                    for (int m = 0; m < M; ++m) u[m] = v[b][m][r];
                    time_op<R, M>(u, w, st, (long long)(j + r * Q) * N2 + n2_0 + c, pk, sm);
#pragma unroll
                    for (int m = 0; m < M; ++m) v[b][m][r] = u[m];
                }
            }
        }
        // the reduction's barriers also order the last-stage LDS reads above
        // before the first-stage writes below
        block_max_sum<TN>(pk, sm, redbuf);
        if (threadIdx.x == 0) {
            Reduction* rd = a.red + (a.counter % 3) * a.n_waves + w_id;
            atomicMax(&rd->peak_bits, double_to_bits(pk));
            atomicAdd(&rd->sum, sm);
            if (w.energy_comp && expected_energy(st) > 0) add_fixed_sum(rd, sm, expected_energy(st));
            // the logarithmic grid's step count depends on the sum; keep the
            // tile sums so that it is formed in tile order, whatever order
            // the workgroups finish in
            if (w.grid == GRID_LOGARITHMIC || (w.energy_comp && !(expected_energy(st) > 0)))
                a.partials[((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride + tile] = sm;
        }
        // the store's twiddle table reads, issued before the forward DFT so
        // that their latency passes during its LDS stages
        cx<double> it0[PER][2], iq1[PER][2];
        if (forward && kRecompute) {
#pragma unroll
            for (int b = 0; b < PER; ++b) {
                const int id = threadIdx.x + b * TN;
                if (ITEMS % TN == 0 || id < ITEMS)
                    inter_loads<R>(tb, (unsigned)(n2_0 + id % C), (unsigned)(id / C), (unsigned)Q, a.log_n, it0[b],
                                   iq1[b]);
            }
        }
        if (forward) {
#pragma unroll
            for (int b = 0; b < PER; ++b) {
                const int id = threadIdx.x + b * TN;
                if (ITEMS % TN == 0 || id < ITEMS) {
                    const int c = id % C, j = id / C;
#pragma unroll
                    for (int m = 0; m < M; ++m) first_stage_store<R, N1, S, -1>(buf, v[b][m], m * C + c, j);
                }
            }
            __syncthreads();
            middle_stages<R, N1, S, TN, -1, true>(buf, tw_t);
        }
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
                const int n2 = n2_0 + c;
                const long long base = at(j, n2);
                if (forward && kRecompute) inter_form<R>(it0[b][0] * it0[b][1], iq1[b][0] * iq1[b][1], itw[b]);
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    if (forward) {
                        last_stage_load<R, N1, S, -1>(buf, tw_t, v[b][m], m * C + c, j);
#pragma unroll
                        for (int r = 0; r < 8; ++r) v[b][m][r] = v[b][m][r] * itw[b][r];
                    }
#pragma unroll
                    for (int r = 0; r < 8; ++r) g[(long long)m * N + base + r * rstride] = v[b][m][r];
                }
            }
        }
        if (t + 1 < TPB) {
            // this tile's last-stage LDS reads before the next tile's writes
            __syncthreads();
#pragma unroll
            for (int b = 0; b < PER; ++b)
#pragma unroll
                for (int m = 0; m < M; ++m)
#pragma unroll
                    for (int r = 0; r < 8; ++r) v[b][m][r] = vn[b][m][r];
        }
    }
}

// Dispersion, loss and gain for the 8 bins k = k1 + N1 (j + r Q) of one work
// item, multiplied into v.  The signed bins k_s form the progression
// K0 + t D, t < 8, D = N/8, K0 = k1 + N1 j - 4 D (r = (t + 4) mod 8), since
// bins above N/2 are k - N; the Nyquist bin's sign does not matter in k^2.
// FP64: exp(i theta k_s^2) by the recurrence f_{t+1} = f_t g_t,
// g_{t+1} = g_t exp(2 i theta D^2) (3 sincos per 8 bins, error ~ 30 ulp).
// FP32: each phase reduced mod 2 pi in FP64, then FP32 sin/cos (the FP64
// recurrence rounded per bin needs 133 instead of 127 registers in the
// two-tile FP32 spectral pass, one wave per SIMD fewer, and measured slower).
template <typename R>
__device__ __forceinline__ void spectral_multiply(cx<R>* v, const WaveStatic& w, int k1, int j, int N1, int Q,
                                                  double arg, double mag) {
    const long long N = w.samples;
    const long long D = N / 8;
    const long long K0 = (long long)k1 + (long long)N1 * j - 4 * D;
    if constexpr (sizeof(R) == 8) {
        const double th = arg * w.dw * w.dw;
        double s, c;
        const double dK0 = (double)K0, dD = (double)D;
        sincos_fast(th * dK0 * dK0, &s, &c);
        cx<double> f = mk<double>(c * mag, s * mag);
        sincos_fast(th * dD * (2.0 * dK0 + dD), &s, &c);
        cx<double> gg = mk<double>(c, s);
        sincos_fast(2.0 * th * dD * dD, &s, &c);
        const cx<double> cc = mk<double>(c, s);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = (t + 4) & 7;
            v[r] = v[r] * f;
            f = f * gg;
            gg = gg * cc;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int t = (r + 4) & 7;
            const double om = (double)(K0 + t * D) * w.dw;
            v[r] = v[r] * phase_factor<R>(arg * om * om, mag);
        }
    }
}

// spectral pass: S rows of length N2; row q = m * N1 + k1 is contiguous.
// A workgroup runs TPB consecutive tiles of one waveform, the next tile's
// loads issued before the present tile's DFTs.
template <typename R, int N2, int S, int TPB>
__global__ __launch_bounds__(pass_threads(S * N2 / 8)) void k_spec_pass(PassArgs a, Tables<R> tb) {
    constexpr int Q = N2 / 8;
    constexpr int ITEMS = S * Q;
    constexpr int TN = pass_threads(ITEMS);
    constexpr int PER = (ITEMS + TN - 1) / TN;
    static_assert(N2 >= 64, "split-stage DFT needs length >= 64");
    using TW = PassTw<R, N2>;
    __shared__ cx<R> buf[padded(N2 * S)];
    __shared__ cx<R> tw_a[TW::A];
    __shared__ cx<R> tw_b[TW::B];
    // plain block order: consecutive tiles are consecutive rows in memory,
    // and spreading them over the XCDs streams best (an XCD-grouped order
    // measured 4.4 against 5.6 TB/s for the same row traffic)
    const int blk = blockIdx.x;
    const int groups = a.tiles_per_wave / TPB;
    const int w_id = a.waves[blk / groups];
    const int tile0 = (blk % groups) * TPB;
    WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    const int was_done = st.phase == PH_DONE;
    // the first spectral pass (after the entry op) reads natural order
    const bool il_in = a.layout_g == 2 && st.top != TOP_ENTRY;
    const bool il_out = a.layout_g == 2;
    const WaveStatic w = a.statics[w_id];
    const Reduction rd = a.red[(a.counter % 3) * a.n_waves + w_id];
    const double* parts = a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride;
    cx<R>* g0 = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    // offset of element j + r Q of row s of the tile: s N2 + j + r Q in
    // natural order, or Interleave(s, j) + 2 r Q (Q a multiple of W)
    auto row_at = [&](bool il, int s, int j) -> int {
        return il ? (int)Interleave<R>::at(s, j, N2) : s * N2 + j;
    };
    auto load = [&](int tile, cx<R> (&dst)[PER][8]) {
        const cx<R>* g = g0 + (long long)tile * S * N2;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int o = row_at(il_in, id % S, id / S), rs = il_in ? 2 * Q : Q;
#pragma unroll
                for (int r = 0; r < 8; ++r) dst[b][r] = g[o + r * rs];
            }
        }
    };
    static_assert(TPB == 1 || TPB == 2, "one or two tiles per workgroup");
    cx<R> va[PER][8], vb[PER][8];
    // the stage-twiddle table's loads and the first tile's are issued before
    // the step is planned, so that their latency overlaps the plan (a
    // finished program's workgroups load nothing)
    const auto tw_staged = TW::template fetch<TN>(tb.bank);
    if (!was_done) load(tile0, va);
    double sum = rd.sum;
    if (w.grid == GRID_LOGARITHMIC && st.phase == PH_SPAN_START) {
        sum = 0;
        for (int t = 0; t < a.time_tiles; ++t) sum += parts[t];
    } else if (w.energy_comp) {
        sum = comp_sum(st, rd, [&] { return tile_sum(parts, a.time_tiles); });
    }
    SpectralOp op = plan_spectral(w, st, bits_to_double(rd.peak_bits), sum);
    if (tile0 == 0 && threadIdx.x == 0) {
        record_step(w, st);
        a.states[((a.counter + 1) & 1) * a.n_waves + w_id] = st;
        Reduction z;
        z.peak_bits = 0;
        z.sum = 0;
        z.fix = 0;
        a.red[((a.counter + 1) % 3) * a.n_waves + w_id] = z;
        if (!was_done && st.phase == PH_DONE) atomicAdd(a.done_count, 1);
    }
    if (!op.active) return;
    const int N = w.samples;
    const int N1 = N / N2;
    const auto tw_t = TW::template store<TN>(tw_a, tw_b, tw_staged);
    // one tile; v holds its values (loaded before the plan for tile 0), vn
    // receives the next tile's
    auto run_tile = [&](auto tc, cx<R> (&v)[PER][8], cx<R> (&vn)[PER][8]) {
        constexpr int t = decltype(tc)::value;
        const int tile = tile0 + t;
        cx<R>* g = g0 + (long long)tile * S * N2;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) first_stage_store<R, N2, S, -1>(buf, v[b], id % S, id / S);
        }
        if constexpr (t + 1 < TPB) load(tile + 1, vn);
        __syncthreads();
        middle_stages<R, N2, S, TN, -1, false>(buf, tw_t);
        const double arg = 0.5 * w.beta2 * op.s;
        const double mag = op.gain * op.loss / (double)N;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int s = id % S, j = id / S;
                last_stage_load<R, N2, S, -1>(buf, tw_t, v[b], s, j);
                spectral_multiply<R>(v[b], w, (tile * S + s) % N1, j, N1, Q, arg, mag);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) first_stage_store<R, N2, S, +1>(buf, v[b], id % S, id / S);
        }
        __syncthreads();
        middle_stages<R, N2, S, TN, +1, true>(buf, tw_t);
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int s = id % S, j = id / S;
                last_stage_load<R, N2, S, +1>(buf, tw_t, v[b], s, j);
                const int o = row_at(il_out, s, j), rs = il_out ? 2 * Q : Q;
#pragma unroll
                for (int r = 0; r < 8; ++r) g[o + r * rs] = v[b][r];
            }
        }
        // this tile's last-stage LDS reads before the next tile's writes
        if constexpr (t + 1 < TPB) __syncthreads();
    };
    run_tile(std::integral_constant<int, 0>{}, va, vb);
    if constexpr (TPB > 1) run_tile(std::integral_constant<int, 1>{}, vb, va);
}

// Row-pair interleaved order back to natural order, in place: one workgroup
// per block of two rows (2 N2 values, contiguous in both orders) of one
// mode of one waveform, staged through LDS (2 N2 values of dynamic shared
// memory).  a.tiles_per_wave = row pairs per waveform (modes N1 / 2).
template <typename R>
__global__ __launch_bounds__(256) void k_deinterleave(PassArgs a, int n2) {
    extern __shared__ unsigned char lds_raw[];
    cx<R>* t = reinterpret_cast<cx<R>*>(lds_raw);
    constexpr int W = Interleave<R>::W;
    const int w_id = a.waves[blockIdx.x / a.tiles_per_wave];
    const int pair = blockIdx.x % a.tiles_per_wave;
    const WaveStatic w = a.statics[w_id];
    cx<R>* g = reinterpret_cast<cx<R>*>(a.field) + w.offset + (long long)pair * 2 * n2;
    for (int i = threadIdx.x; i < 2 * n2; i += blockDim.x) {
        const int within = i % (2 * W);
        t[(within / W) * n2 + (i / (2 * W)) * W + within % W] = g[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * n2; i += blockDim.x) g[i] = t[i];
}

}  // namespace focss_hip
