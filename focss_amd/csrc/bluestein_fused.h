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
// once, after the program's last time op.  Both kernels use the split-stage
// DFTs of the SSFM passes (first and last radix-8 stage in registers, one
// LDS round trip per middle stage).
//
// UNIFORM programs, whose spectral ops apply at most two distinct step
// lengths s, skip the chirps: the spectral op is the length-N circular
// convolution with d_s = IDFT_N(D_s), which equals the length-M circular
// convolution of the zero-padded field with K_s = [d_0 .. d_{N-1}, 0 .. 0,
// d_1 .. d_{N-1}] on its first N values (M >= 2N - 1, so the wrapped
// terms never overlap).  With Khat_s = DFT_M(K_s) / M tabulated in row
// order, a split step is two launches:
//   spectral op    k_blu_rows<PAD>: plan (as the spectral pass), forward
//                  row DFT -> x Khat_s x gain e^{-alpha s / 2} -> inverse
//                  row DFT
//   time op        k_blu_cols<PAD_OUT>: column IDFT -> the op on e < N,
//                  zero beyond -> column DFT (PAD_ENTRY at entry)
// and about 5 M values move per split step.
#pragma once
#include "bluestein_kernels.h"
#include "fft_lds.h"
#include "ssfm_kernels.h"

namespace focss_hip {

enum { BLU_ENTRY = 0, BLU_MID = 1, BLU_OUT = 2, BLU_PAD_ENTRY = 3, BLU_PAD_OUT = 4 };

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

// the split of M: N1 = 2^max(6, floor((log_m - 1) / 2)) (N1 <= N2, both at
// least 64 for the split-stage DFTs)
constexpr int blu_log_n1(int log_m) { return (log_m - 1) / 2 < 6 ? 6 : (log_m - 1) / 2; }
template <typename R, int LOGM>
struct BluShape {
    static constexpr int N1 = 1 << blu_log_n1(LOGM);
    static constexpr int N2 = (1 << LOGM) / N1;
    // values per workgroup: rows 1024 (complex128) / 2048 (complex64), columns 2048
    static constexpr int ROW_ELEMS = sizeof(R) == 8 ? 1024 : 2048;
    static constexpr int S = ROW_ELEMS / N2 < 1 ? 1 : (ROW_ELEMS / N2 > N1 ? N1 : ROW_ELEMS / N2);
    template <int MODES>
    static constexpr int C = 2048 / (N1 * MODES) < 1 ? 1 : (2048 / (N1 * MODES) > N2 ? N2 : 2048 / (N1 * MODES));
};

// S rows (k1 = k1_0 + s) of one mode of one waveform, split-stage as the
// spectral pass: forward row DFT N2 -> x Bhat (row order) -> inverse row
// DFT N2 -> x W_M^{-n2 k1}.  after_mid: the second convolution of the
// spectral op (runs when the op is active), else the first (runs while the
// program has not ended).  PAD: the whole spectral op of a UNIFORM program,
// planned here, with Khat_s and the op's scalar in place of Bhat.  A work
// item (s, j < N2/8) owns the values j + r N2/8 of row s.
template <typename R, int LOGM, bool PAD>
__global__ __launch_bounds__(pass_threads(BluShape<R, LOGM>::S * BluShape<R, LOGM>::N2 / 8)) void k_blu_rows(
    BluArgs<R> g, int after_mid, Tables<R> tb) {
    using Sh = BluShape<R, LOGM>;
    constexpr int N1 = Sh::N1, N2 = Sh::N2, S = Sh::S, M = 1 << LOGM;
    constexpr int Q = N2 / 8;
    constexpr int ITEMS = S * Q;
    constexpr int TN = pass_threads(ITEMS);
    constexpr int PER = (ITEMS + TN - 1) / TN;
    using TW = PassTw<R, N2>;
    __shared__ cx<R> buf[padded(N2 * S)];
    __shared__ cx<R> tw_a[TW::A];
    __shared__ cx<R> tw_b[TW::B];
    const int k = blockIdx.x / g.tiles, t = blockIdx.x % g.tiles;
    const int m = t / (N1 / S);
    const int k1_0 = (t % (N1 / S)) * S;
    const int w_id = g.p.waves[k];
    const cx<R>* bh;
    R kscale = 1;
    if (PAD) {
        const PassArgs& a = g.p;
        WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
        const WaveStatic w = a.statics[w_id];
        const int was_done = st.phase == PH_DONE;
        const Reduction rd = a.red[(a.counter % 3) * a.n_waves + w_id];
        // the column kernel's tile sums in tile order (UNIFORM programs use
        // the sum for energy_comp only)
        const double sum = !w.energy_comp ? rd.sum : comp_sum(st, rd, [&] {
            return tile_sum(a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride,
                            N2 / (w.modes == 1 ? Sh::template C<1> : Sh::template C<2>));
        });
        const SpectralOp op = plan_spectral(w, st, bits_to_double(rd.peak_bits), sum);
        if (t == 0 && threadIdx.x == 0) {
            a.states[((a.counter + 1) & 1) * a.n_waves + w_id] = st;
            Reduction z;
            z.peak_bits = 0;
            z.sum = 0;
            z.fix = 0;
            a.red[((a.counter + 1) % 3) * a.n_waves + w_id] = z;
            if (!was_done && st.phase == PH_DONE) atomicAdd(a.done_count, 1);
        }
        if (!op.active) return;
        // the table of this step length (the host tabulated every length the
        // program applies)
        const GenWave<R>& gw = g.gw[w_id];
        const int sel = fabs(op.s - gw.kval[1]) < fabs(op.s - gw.kval[0]);
        bh = gw.khat[sel] + (long long)k1_0 * N2;
        kscale = (R)(op.gain * op.loss);
    } else {
        if (after_mid) {
            if (!g.active[w_id]) return;
        } else if (g.p.states[(g.p.counter & 1) * g.p.n_waves + w_id].phase == PH_DONE) {
            return;
        }
        bh = g.gw[w_id].bhat_rows + (long long)k1_0 * N2;
    }
    const int modes = g.p.statics[w_id].modes;
    cx<R>* rows = g.work + ((long long)k * modes + m) * M + (long long)k1_0 * N2;
    const auto tw = TW::template load<TN>(tw_a, tw_b, tb.bank);
    cx<R> v[PER][8];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) {
            const int s = id % S, j = id / S;
#pragma unroll
            for (int r = 0; r < 8; ++r) v[b][r] = rows[s * N2 + j + r * Q];
            first_stage_store<R, N2, S, -1>(buf, v[b], s, j);
        }
    }
    __syncthreads();
    middle_stages<R, N2, S, TN, -1, false>(buf, tw);
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) {
            const int s = id % S, j = id / S;
            last_stage_load<R, N2, S, -1>(buf, tw, v[b], s, j);
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const cx<R> f = bh[s * N2 + j + r * Q];
                v[b][r] = v[b][r] * (PAD ? f * kscale : f);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) first_stage_store<R, N2, S, +1>(buf, v[b], id % S, id / S);
    }
    __syncthreads();
    middle_stages<R, N2, S, TN, +1, true>(buf, tw);
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) {
            const int s = id % S, j = id / S;
            last_stage_load<R, N2, S, +1>(buf, tw, v[b], s, j);
            cx<R> itw[8];  // W_M^{k1 (j + r Q)}
            inter_twiddles<R>(tb, (unsigned)(k1_0 + s), (unsigned)j, (unsigned)Q, LOGM, itw);
#pragma unroll
            for (int r = 0; r < 8; ++r) rows[s * N2 + j + r * Q] = v[b][r] * conj(itw[r]);
        }
    }
}

// C columns (all modes) of one waveform, split-stage as the time pass:
// inverse column DFT N1 (not at ENTRY) -> the pointwise operator of KIND on
// the time-domain values e = n1 N2 + n2 (see the header comment; zero
// beyond N) -> forward column DFT N1 -> x W_M^{n2 k1}; after the program's
// last op the field is written instead.  A work item (c, j < N1/8) owns the
// values j + r N1/8 of column c of every mode.
template <typename R, int LOGM, int MODES, int KIND>
__global__ __launch_bounds__(pass_threads(BluShape<R, LOGM>::template C<MODES> * BluShape<R, LOGM>::N1 / 8)) void
k_blu_cols(BluArgs<R> g, Tables<R> tb) {
    using Sh = BluShape<R, LOGM>;
    constexpr int N1 = Sh::N1, N2 = Sh::N2, C = Sh::template C<MODES>, M = 1 << LOGM;
    constexpr int SEQ = C * MODES;  // sequence (m, c) at LDS column m C + c
    constexpr int Q = N1 / 8;
    constexpr int ITEMS = C * Q;
    constexpr int TN = pass_threads(ITEMS);
    constexpr int PER = (ITEMS + TN - 1) / TN;
    using TW = PassTw<R, N1>;
    __shared__ cx<R> buf[padded(N1 * SEQ)];
    __shared__ cx<R> tw_a[TW::A];
    __shared__ cx<R> tw_b[TW::B];
    __shared__ double redbuf[2 * (TN / 64)];
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
        } else if (w.energy_comp) {
            sum = comp_sum(st, rd, [&] {
                return tile_sum(a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride,
                                g.tiles);
            });
        }
        op = plan_spectral(w, st, bits_to_double(rd.peak_bits), sum);
        if (t == 0 && threadIdx.x == 0) {
            record_step(w, st);
            a.states[((a.counter + 1) & 1) * a.n_waves + w_id] = st;
            Reduction z;
            z.peak_bits = 0;
            z.sum = 0;
            z.fix = 0;
            a.red[((a.counter + 1) % 3) * a.n_waves + w_id] = z;
            g.active[w_id] = op.active;
            if (!was_done && st.phase == PH_DONE) atomicAdd(a.done_count, 1);
        }
        if (!op.active) return;
    } else if (st.top == TOP_NONE) {
        return;
    }
    constexpr bool PAD = KIND == BLU_PAD_ENTRY || KIND == BLU_PAD_OUT;
    constexpr bool ENTRY = KIND == BLU_ENTRY || KIND == BLU_PAD_ENTRY;
    const GenWave<R> gw = g.gw[w_id];
    const int n = gw.n;
    cx<R>* rows = g.work + (long long)k * MODES * M;
    cx<R>* x = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    const auto tw = TW::template load<TN>(tw_a, tw_b, tb.bank);
    cx<R> v[PER][MODES][8];
    if (ENTRY) {
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int mm = 0; mm < MODES; ++mm)
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        const int e = (j + r * Q) * N2 + n2_0 + c;
                        v[b][mm][r] = e < n ? x[(long long)mm * n + e] : mk<R>(0, 0);
                    }
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int mm = 0; mm < MODES; ++mm) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) v[b][mm][r] = rows[(long long)mm * M + (j + r * Q) * N2 + n2_0 + c];
                    first_stage_store<R, N1, SEQ, +1>(buf, v[b][mm], mm * C + c, j);
                }
            }
        }
        __syncthreads();
        middle_stages<R, N1, SEQ, TN, +1, false>(buf, tw);
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int mm = 0; mm < MODES; ++mm) last_stage_load<R, N1, SEQ, +1>(buf, tw, v[b][mm], mm * C + c, j);
            }
        }
    }
    bool last = false;
    if (KIND == BLU_MID) {
        const double arg = 0.5 * w.beta2 * op.s;
        const double mag = op.gain * op.loss / (double)n;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int e = (j + r * Q) * N2 + n2_0 + c;
                    if (e < n) {
                        const cx<R> ce = gw.chirp[e];
                        const cx<R> f = spectral_factor<R>(w, op, e, arg, mag);
#pragma unroll
                        for (int mm = 0; mm < MODES; ++mm) v[b][mm][r] = conj((ce * v[b][mm][r]) * f) * ce;
                    } else {
#pragma unroll
                        for (int mm = 0; mm < MODES; ++mm) v[b][mm][r] = mk<R>(0, 0);
                    }
                }
            }
        }
    } else {
        // the time-domain op on all modes of one sample at a time
        last = st.phase == PH_DONE;  // this op is the program's last
        double pk = 0, sm = 0;
#pragma unroll
        for (int b = 0; b < PER; ++b) {
            const int id = threadIdx.x + b * TN;
            if (ITEMS % TN == 0 || id < ITEMS) {
                const int c = id % C, j = id / C;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int e = (j + r * Q) * N2 + n2_0 + c;
                    if (e < n) {
                        const cx<R> ce = PAD ? mk<R>(1, 0) : gw.chirp[e];
                        cx<R> u[MODES];
#pragma unroll
                        for (int mm = 0; mm < MODES; ++mm)
                            u[mm] = ENTRY || PAD ? v[b][mm][r] : conj(ce * v[b][mm][r]);
                        time_op<R, MODES>(u, w, st, e, pk, sm);
#pragma unroll
                        for (int mm = 0; mm < MODES; ++mm) {
                            if (last) x[(long long)mm * n + e] = u[mm];
                            v[b][mm][r] = PAD ? u[mm] : u[mm] * ce;
                        }
                    } else {
#pragma unroll
                        for (int mm = 0; mm < MODES; ++mm) v[b][mm][r] = mk<R>(0, 0);
                    }
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
            if (w.grid == GRID_LOGARITHMIC || (w.energy_comp && !(expected_energy(st) > 0)))
                a.partials[((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride + t] = sm;
        }
        if (last) return;
    }
    if (KIND == BLU_MID) __syncthreads();  // the last-stage LDS reads before the first-stage writes
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) {
            const int c = id % C, j = id / C;
#pragma unroll
            for (int mm = 0; mm < MODES; ++mm) first_stage_store<R, N1, SEQ, -1>(buf, v[b][mm], mm * C + c, j);
        }
    }
    __syncthreads();
    middle_stages<R, N1, SEQ, TN, -1, true>(buf, tw);
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = threadIdx.x + b * TN;
        if (ITEMS % TN == 0 || id < ITEMS) {
            const int c = id % C, j = id / C;
            cx<R> itw[8];  // W_M^{n2 (j + r Q)}
            inter_twiddles<R>(tb, (unsigned)(n2_0 + c), (unsigned)j, (unsigned)Q, LOGM, itw);
#pragma unroll
            for (int mm = 0; mm < MODES; ++mm) {
                last_stage_load<R, N1, SEQ, -1>(buf, tw, v[b][mm], mm * C + c, j);
#pragma unroll
                for (int r = 0; r < 8; ++r) rows[(long long)mm * M + (j + r * Q) * N2 + n2_0 + c] = v[b][mm][r] * itw[r];
            }
        }
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

// Building Khat_s for a UNIFORM program (FP64): D_s(k) = exp(i arg w_k^2),
// arg = beta2 s / 2, w_k = kk dw with kk in (-N/2, N/2] (the spectral
// factor of the passes without loss and gain)
static __global__ void k_pad_disp(cx<double>* __restrict__ d, int n, double arg, double dw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int kk = k <= n / 2 ? k : k - n;
    const double om = (double)kk * dw;
    double sn, cs;
    sincos(arg * (om * om), &sn, &cs);
    d[k] = mk<double>(cs, sn);
}

// K (M values) from the unscaled inverse DFT u = N d of D_s:
// K_j = d_j (j < N), K_{M-j} = d_{N-j} (0 < j < N), 0 between
static __global__ void k_pad_kernel(const cx<double>* __restrict__ u, cx<double>* __restrict__ kv, int n, int log_m) {
    const int M = 1 << log_m;
    const double inv_n = 1.0 / n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        cx<double> v = mk<double>(0, 0);
        if (i < n)
            v = u[i] * inv_n;
        else if (i > M - n)
            v = u[i - (M - n)] * inv_n;
        kv[i] = v;
    }
}

// Khat in the row order of the split, at the field's precision
template <typename R>
__global__ void k_pad_rows(const cx<double>* __restrict__ in, cx<R>* __restrict__ out, int log_m, int log_n1) {
    const int M = 1 << log_m, N2 = M >> log_n1, N1 = 1 << log_n1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const cx<double> v = in[(i / N2) + N1 * (i % N2)];
        out[i] = mk<R>((R)v.x, (R)v.y);
    }
}

}  // namespace focss_hip
