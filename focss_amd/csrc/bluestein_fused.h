// Link programs of generic length N with Bluestein length M = 2^log_m >=
// 2^12, in four launches per split step.  The length-M DFTs are split
// M = N1 N2 (index n = n1 N2 + n2 in time, k = k1 + N1 k2 in frequency) as
//     forward: column DFT_N1 over n1 -> k1, x W_M^{n2 k1}, row DFT_N2 over n2 -> k2
//     inverse: row IDFT_N2 over k2 -> n2, x W_M^{-n2 k1}, column IDFT_N1 over k1 -> n1
// and the frequency domain stays in row order (k1 N2 + k2), where Bhat is
// stored in the same order, so one row kernel does forward row DFT, x Bhat,
// inverse row DFT.  Each column kernel ends one convolution and starts the
// next, with the pointwise operator between them on the time-domain values
// of its columns:
//   entry          k_blu_cols<ENTRY>: field -> time op (span scale, power
//                  reductions) -> x c -> column DFT
//   spectral op    k_blu_rows -> k_blu_cols<MID>: column IDFT -> X = c conv,
//                  plan, Y = X D / N, conj(Y) c -> column DFT -> k_blu_rows
//   time op        k_blu_cols<OUT>: column IDFT -> x = conj(c conv) -> time
//                  op -> x c -> column DFT (or, after the last op, x to the
//                  field)
// Between launches only the work rows (modes x M values per waveform) move
// through HBM: about 10 M values per split step, against 2 N for the
// power-of-two two-pass programs.  The field is read at entry and written
// once, after the program's last time op.
#pragma once
#include "bluestein_kernels.h"
#include "fft_lds.h"

namespace focss_hip {

enum { BLU_ENTRY = 0, BLU_MID = 1, BLU_OUT = 2 };

constexpr int kBluMinLog = 12;  // smaller Bluestein lengths run the k_gen_* kernels

template <typename R>
struct BluArgs {
    PassArgs p;            // field, control arrays, waves of this class, counter
    const GenWave<R>* gw;  // indexed by wave id
    cx<R>* work;           // class rows: wave k, mode m at (k modes + m) M
    int* active;           // per wave id: the spectral op of this counter runs
    int log_m;
    int tiles;             // workgroups per waveform of this launch
};

// the split of M: N1 = 2^floor((log_m - 1) / 2) (N1 <= N2, columns of at
// least 4 values per workgroup at every length)
constexpr int blu_log_n1(int log_m) { return (log_m - 1) / 2; }
template <typename R>
constexpr int blu_tile() { return sizeof(R) == 8 ? 4096 : 8192; }  // values per workgroup (64 KiB)
template <typename R, int LOGM>
struct BluShape {
    static constexpr int N1 = 1 << blu_log_n1(LOGM);
    static constexpr int N2 = (1 << LOGM) / N1;
    static constexpr int S = (blu_tile<R>() / N2 < N1) ? blu_tile<R>() / N2 : N1;  // rows per row-kernel tile
    template <int MODES>
    static constexpr int C = (blu_tile<R>() / (N1 * MODES) < N2) ? blu_tile<R>() / (N1 * MODES) : N2;
};

// forward row DFT, x Bhat (row order), inverse row DFT, x W_M^{-n2 k1}, on S
// rows of one mode of one waveform; after_mid: the second convolution of
// the spectral op (runs when the op is active), else the first (runs while
// the program has not ended)
template <typename R, int LOGM>
__global__ __launch_bounds__(NT) void k_blu_rows(BluArgs<R> g, int after_mid, Tables<R> tb) {
    using Sh = BluShape<R, LOGM>;
    constexpr int N1 = Sh::N1, N2 = Sh::N2, S = Sh::S, M = 1 << LOGM;
    constexpr int LB = (ilog2<N2>::value + 1) / 2;
    __shared__ cx<R> buf[padded(N2 * S)];
    __shared__ cx<R> tw_lo[1 << LB];
    __shared__ cx<R> tw_hi[N2 >> LB];
    const int k = blockIdx.x / g.tiles, t = blockIdx.x % g.tiles;
    const int m = t / (N1 / S);
    const int k1_0 = (t % (N1 / S)) * S;
    const int w_id = g.p.waves[k];
    if (after_mid) {
        if (!g.active[w_id]) return;
    } else if (g.p.states[(g.p.counter & 1) * g.p.n_waves + w_id].phase == PH_DONE) {
        return;
    }
    const GenWave<R> gw = g.gw[w_id];
    const int modes = g.p.statics[w_id].modes;
    cx<R>* rows = g.work + ((long long)k * modes + m) * M + (long long)k1_0 * N2;
    const cx<R>* bh = gw.bhat_rows + (long long)k1_0 * N2;
    load_tw_split<R, N2, LB, NT>(tw_lo, tw_hi, tb.bank);
    const TwSplit<R, LB> tw{tw_lo, tw_hi};
    for (int i = threadIdx.x; i < N2 * S; i += NT) buf[pad((i % N2) * S + i / N2)] = rows[i];
    __syncthreads();
    fft_lds<R, N2, S, NT, -1>(buf, tw);
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        const int p = pad((i % N2) * S + i / N2);
        buf[p] = buf[p] * bh[i];
    }
    __syncthreads();
    fft_lds<R, N2, S, NT, +1>(buf, tw);
    for (int i = threadIdx.x; i < N2 * S; i += NT) {
        const int s = i / N2, n2 = i % N2;
        const cx<R> w = tw_big<R>(tb, (unsigned)(n2 * (k1_0 + s)) & (unsigned)(M - 1), LOGM);
        rows[i] = buf[pad(n2 * S + s)] * conj(w);
    }
}

// C columns (all modes) of one waveform: see the header comment for KIND
template <typename R, int LOGM, int MODES, int KIND>
__global__ __launch_bounds__(NT) void k_blu_cols(BluArgs<R> g, Tables<R> tb) {
    using Sh = BluShape<R, LOGM>;
    constexpr int N1 = Sh::N1, N2 = Sh::N2, C = Sh::template C<MODES>, M = 1 << LOGM;
    constexpr int SEQ = C * MODES;  // sequence (m, c) at LDS column m C + c
    constexpr int LB = (ilog2<N1>::value + 1) / 2;
    __shared__ cx<R> buf[padded(N1 * SEQ)];
    __shared__ cx<R> tw_lo[1 << LB];
    __shared__ cx<R> tw_hi[N1 >> LB];
    __shared__ double redbuf[2 * (NT / 64)];
    const PassArgs& a = g.p;
    const int k = blockIdx.x / g.tiles, t = blockIdx.x % g.tiles;
    const int n2_0 = t * C;
    const int w_id = a.waves[k];
    WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    const WaveStatic w = a.statics[w_id];
    SpectralOp op;
    if (KIND == BLU_MID) {
        const int was_done = st.phase == PH_DONE;
        const Reduction rd = a.red[(a.counter % 3) * a.n_waves + w_id];
        double sum = rd.sum;
        if (w.grid == GRID_LOGARITHMIC && st.phase == PH_SPAN_START) {
            // the time op's tile sums in tile order (run-to-run deterministic)
            const double* p = a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride;
            sum = 0;
            for (int i = 0; i < g.tiles; ++i) sum += p[i];
        }
        op = plan_spectral(w, st, bits_to_double(rd.peak_bits), sum);
        if (t == 0 && threadIdx.x == 0) {
            a.states[((a.counter + 1) & 1) * a.n_waves + w_id] = st;
            Reduction z;
            z.peak_bits = 0;
            z.sum = 0;
            a.red[((a.counter + 1) % 3) * a.n_waves + w_id] = z;
            g.active[w_id] = op.active;
            if (!was_done && st.phase == PH_DONE) atomicAdd(a.done_count, 1);
        }
        if (!op.active) return;
    } else if (st.top == TOP_NONE) {
        return;
    }
    const GenWave<R> gw = g.gw[w_id];
    const int n = gw.n;
    cx<R>* rows = g.work + (long long)k * MODES * M;
    cx<R>* x = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    load_tw_split<R, N1, LB, NT>(tw_lo, tw_hi, tb.bank);
    const TwSplit<R, LB> tw{tw_lo, tw_hi};
    if (KIND == BLU_ENTRY) {
        for (int i = threadIdx.x; i < N1 * SEQ; i += NT) {
            const int n1 = i / SEQ, q = i % SEQ, mm = q / C, e = n1 * N2 + n2_0 + q % C;
            buf[pad(i)] = e < n ? x[(long long)mm * n + e] : mk<R>(0, 0);
        }
    } else {
        for (int i = threadIdx.x; i < N1 * SEQ; i += NT) {
            const int k1 = i / SEQ, q = i % SEQ, mm = q / C;
            buf[pad(i)] = rows[(long long)mm * M + (long long)k1 * N2 + n2_0 + q % C];
        }
        __syncthreads();
        fft_lds<R, N1, SEQ, NT, +1>(buf, tw);  // conv at n1 N2 + n2
    }
    __syncthreads();
    bool last = false;
    if (KIND == BLU_MID) {
        const double arg = 0.5 * w.beta2 * op.s;
        const double mag = op.gain * exp(-0.5 * w.alpha * op.s) / (double)n;
        for (int i = threadIdx.x; i < N1 * SEQ; i += NT) {
            const int e = (i / SEQ) * N2 + n2_0 + (i % SEQ) % C;
            cx<R> v = mk<R>(0, 0);
            if (e < n) {
                const cx<R> c = gw.chirp[e];
                v = conj((c * buf[pad(i)]) * spectral_factor<R>(w, op, e, arg, mag)) * c;
            }
            buf[pad(i)] = v;
        }
    } else {
        // time-domain values of all modes of one sample per thread
        last = st.phase == PH_DONE;  // this op is the program's last
        double pk = 0, sm = 0;
        for (int i = threadIdx.x; i < N1 * C; i += NT) {
            const int n1 = i / C, cc = i % C, e = n1 * N2 + n2_0 + cc;
            cx<R> u[MODES];
            cx<R> c = mk<R>(0, 0);
            if (e < n) {
                c = gw.chirp[e];
#pragma unroll
                for (int mm = 0; mm < MODES; ++mm) {
                    const cx<R> v = buf[pad(n1 * SEQ + mm * C + cc)];
                    u[mm] = KIND == BLU_ENTRY ? v : conj(c * v);
                }
                time_op<R, MODES>(u, w, st, e, pk, sm);
            }
#pragma unroll
            for (int mm = 0; mm < MODES; ++mm) {
                if (last) {
                    if (e < n) x[(long long)mm * n + e] = u[mm];
                } else {
                    buf[pad(n1 * SEQ + mm * C + cc)] = e < n ? u[mm] * c : mk<R>(0, 0);
                }
            }
        }
        block_max_sum<NT>(pk, sm, redbuf);
        if (threadIdx.x == 0) {
            Reduction* r = a.red + (a.counter % 3) * a.n_waves + w_id;
            atomicMax(&r->peak_bits, double_to_bits(pk));
            atomicAdd(&r->sum, sm);
            if (w.grid == GRID_LOGARITHMIC)
                a.partials[((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride + t] = sm;
        }
        if (last) return;
    }
    __syncthreads();
    fft_lds<R, N1, SEQ, NT, -1>(buf, tw);
    for (int i = threadIdx.x; i < N1 * SEQ; i += NT) {
        const int k1 = i / SEQ, q = i % SEQ, mm = q / C, n2 = n2_0 + q % C;
        const cx<R> wv = tw_big<R>(tb, (unsigned)(n2 * k1) & (unsigned)(M - 1), LOGM);
        rows[(long long)mm * M + (long long)k1 * N2 + n2] = buf[pad(i)] * wv;
    }
}

// Bhat in the row order of the split: out[k1 N2 + k2] = in[k1 + N1 k2]
template <typename R>
__global__ void k_blu_row_order(const cx<R>* __restrict__ in, cx<R>* __restrict__ out, int log_m, int log_n1) {
    const int M = 1 << log_m, N2 = M >> log_n1, N1 = 1 << log_n1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const int k1 = i / N2, k2 = i % N2;
        out[i] = in[k1 + N1 * k2];
    }
}

}  // namespace focss_hip
