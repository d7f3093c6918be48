// Learned few-step digital backpropagation (no counterpart in the
// reference; measured against plain DBP).  Step j of J:
//     y(t)   = sum_k h_{j,k} x((t - k + c) mod N),  c = K / 2       (complex FIR)
//     p(t)   = |y(t)|^2
//     phi(t) = sum_m g_{j,m} p((t - m + P) mod N),  P = KP / 2      (filtered intensity)
//     z(t)   = y(t) exp(i phi(t))                                   (nonlinear phase)
// KP = 1 is the plain Kerr step exp(i theta_j |y|^2); a wider KP low-pass
// filters the intensity that drives the phase, which lets a long step
// account for the dispersive walk-off inside it.
// Loss = sum over the batch of (1/(M N B)) sum_{mode, t} |z_J(t) - target(t)|^2.
// Gradients are Wirtinger dL/d(conj h) and real dL/dg:
//     q = -Im(conj(G_z) z),   dL/dphi = 2 q,
//     dL/dg_m = sum_t 2 q(t) p(t - m + P),   dL/dp(s) = sum_m g_m 2 q(s + m - P),
//     G_y = G_z e^{-i phi} + (dL/dp) y,
//     G_h_k = sum_t G_y(t) conj(x(t - k + c)),   G_x(s) = sum_k conj(h_k) G_y(s + k - c).
// With M = 2 modes (Manakov) the FIR acts on each polarisation and p(t) is
// the intensity summed over both, so the phase is common to them; the loss
// averages over every mode and sample, and
//     q = -sum_modes Im(conj(G_z) z),   G_h_k sums over the modes.
// The forward pass runs one workgroup per waveform (M N <= 16384 complex64
// in LDS) through every step; the backward pass runs one launch per step
// over chunks of every waveform (k_ldbp_bwd_step).
#pragma once
#include "cx.h"
#include "trig.h"

namespace focss_hip {

constexpr int LDBP_MAX_N = 16384;
constexpr int LDBP_MIN_N = 1024;
constexpr int LDBP_R = 16;          // forward pass: consecutive outputs per thread (N / 16 threads)
constexpr int LDBP_KC = 8;          // forward pass: FIR taps per register window

struct LdbpArgs {
    const cx<float>* in;      // [B][M][N]
    cx<float>* out;           // [B][M][N] (may be null)
    const cx<float>* target;  // [B][M][N] (may be null)
    const cx<float>* fir;     // [J][K]
    const float* kerr;        // [J][KP]
    cx<float>* xs;            // [B][J][M][N] FIR inputs (training) or null
    cx<float>* ys;            // [B][J][M][N] FIR outputs (training) or null
    double* loss;             // [1] accumulated, may be null
    double* grad;             // [J*K*2 + J*KP] accumulated: Re/Im of G_h, then dL/dg
    double* part;             // [J][backward groups][2K + KP] per-workgroup gradient sums
    int J, K, KP, N, B;
};

// forward-pass LDS index: one padding element every 16, so the 16-element
// blocks that neighbouring lanes own start in different banks
__host__ __device__ constexpr int p16(int i) { return i + (i >> 4); }

// wave-uniform read through the scalar data cache
__device__ __forceinline__ float sload(const float* p) {
    return *(const __attribute__((address_space(4))) float*)p;
}

__device__ __forceinline__ double ldbp_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double ldbp_block_sum(double v, double* red, int nt) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nt / 64; ++i) s += red[i];
    return s;  // valid in thread 0
}

// Forward pass, N / 16 threads: thread i owns outputs 16 i .. 16 i + 15 of
// every mode.  LDS holds the M modes, mode stride N + N / 16 (p16 layout).
template <int K, int KP, int M>
__global__ __launch_bounds__(LDBP_MAX_N / LDBP_R) void k_ldbp_fwd(LdbpArgs a) {
    constexpr int R = LDBP_R, KC = LDBP_KC, P = KP / 2, c = K / 2;
    constexpr int NW = LDBP_MAX_N / LDBP_R / 64;
    extern __shared__ unsigned char smem_raw[];
    cx<float>* buf = reinterpret_cast<cx<float>*>(smem_raw);  // [M][p16(N)]
    __shared__ float edge[2][NW][P > 0 ? P : 1];              // first / last P intensities of each wave
    __shared__ double red[NW];
    const int N = a.N, mask = N - 1, S = N + N / 16;
    const int nt = blockDim.x, nw = nt >> 6;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int t0 = threadIdx.x * R;
    const cx<float>* in = a.in + (long long)b * M * N;
#pragma unroll
    for (int md = 0; md < M; ++md)
        for (int t = threadIdx.x; t < N; t += nt) buf[md * S + p16(t)] = in[(long long)md * N + t];
    __syncthreads();
    for (int j = 0; j < a.J; ++j) {
        const float* hj = reinterpret_cast<const float*>(a.fir + (long long)j * K);
        if (a.xs) {
#pragma unroll
            for (int md = 0; md < M; ++md) {
                cx<float>* xs = a.xs + (((long long)b * a.J + j) * M + md) * N + t0;
#pragma unroll
                for (int r = 0; r < R; ++r) xs[r] = buf[md * S + p16(t0 + r)];
            }
        }
        // FIR of each mode, KC taps per register window of R + KC - 1 inputs
        cx<float> y[M][R];
#pragma unroll
        for (int md = 0; md < M; ++md) {
#pragma unroll
            for (int r = 0; r < R; ++r) y[md][r] = mk<float>(0, 0);
            const cx<float>* bm = buf + md * S;
#pragma unroll 1
            for (int k0 = 0; k0 < K; k0 += KC) {
                cx<float> w[R + KC - 1];
                const int base = t0 + c - k0 - (KC - 1);
#pragma unroll
                for (int u = 0; u < R + KC - 1; ++u) w[u] = bm[p16((base + u) & mask)];
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const bool live = k0 + kk < K;
                    const float hr = live ? sload(hj + 2 * (k0 + kk)) : 0.0f;
                    const float hi = live ? sload(hj + 2 * (k0 + kk) + 1) : 0.0f;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const cx<float> x = w[r - kk + KC - 1];
                        y[md][r].x = fmaf(hr, x.x, fmaf(-hi, x.y, y[md][r].x));
                        y[md][r].y = fmaf(hr, x.y, fmaf(hi, x.x, y[md][r].y));
                    }
                }
            }
        }
        if (a.ys) {
#pragma unroll
            for (int md = 0; md < M; ++md) {
                cx<float>* ys = a.ys + (((long long)b * a.J + j) * M + md) * N + t0;
#pragma unroll
                for (int r = 0; r < R; ++r) ys[r] = y[md][r];
            }
        }
        float p[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            p[r] = norm2(y[0][r]);
#pragma unroll
            for (int md = 1; md < M; ++md) p[r] += norm2(y[md][r]);
        }
        if constexpr (P > 0) {
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < P; ++i) edge[0][wid][i] = p[i];
            }
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < P; ++i) edge[1][wid][i] = p[R - P + i];
            }
        }
        // every FIR read of buf is done before any output is written; the
        // wave-edge intensities are visible
        __syncthreads();
        float pe[R + 2 * P];  // p(t0 - P + u)
#pragma unroll
        for (int r = 0; r < R; ++r) pe[P + r] = p[r];
        if constexpr (P > 0) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                float prev = __shfl(p[R - P + i], (lane + 63) & 63, 64);
                float next = __shfl(p[i], (lane + 1) & 63, 64);
                if (lane == 0) prev = edge[1][(wid + nw - 1) % nw][i];
                if (lane == 63) next = edge[0][(wid + 1) % nw][i];
                pe[i] = prev;
                pe[P + R + i] = next;
            }
        }
        float g[KP];
#pragma unroll
        for (int m = 0; m < KP; ++m) g[m] = sload(a.kerr + j * KP + m);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float phi = 0;
#pragma unroll
            for (int m = 0; m < KP; ++m) phi = fmaf(g[m], pe[r - m + 2 * P], phi);
            float s, co;
            sincos_fast(phi, &s, &co);
#pragma unroll
            for (int md = 0; md < M; ++md) buf[md * S + p16(t0 + r)] = y[md][r] * mk<float>(co, s);
        }
        __syncthreads();
    }
    double part = 0;
#pragma unroll
    for (int md = 0; md < M; ++md) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const long long t = ((long long)b * M + md) * N + t0 + r;
            const cx<float> z = buf[md * S + p16(t0 + r)];
            if (a.out) a.out[t] = z;
            if (a.target) part += (double)norm2(z - a.target[t]);
        }
    }
    if (a.loss) {
        double s = ldbp_block_sum(part, red, nt);
        if (threadIdx.x == 0) atomicAdd(a.loss, s / ((double)N * M * a.B));
    }
}

// ------------------------------------------------------------------
// Backward pass split over the whole device: one launch per step j (from
// J - 1 down to 0), one workgroup per LDBP_CH-sample chunk of one waveform
// (all modes).  A launch forms the step's output gradient G_z on the chunk
// and a LDBP_HL-sample halo either side -- from (out - target) for the last
// step, else as the input gradient of step j + 1 from the G_y that the
// previous launch left in global memory -- then runs the nonlinear-phase
// backward (q, phi over chunk and halo, G_y over the chunk), writes G_y for
// the next launch and adds the step's tap gradients.  The kernel boundary is
// the only synchronisation between chunks; the halo work (2 LDBP_HL of
// LDBP_CH samples) is done twice.  Every operand window (G_y of step j + 1,
// y and x of step j) is staged into LDS by coalesced loads first; the two
// FIR-shaped contractions run on units of LDBP_RB consecutive samples with
// register windows of LDBP_KB taps.
// ------------------------------------------------------------------
constexpr int LDBP_CH = 512;   // samples per workgroup
constexpr int LDBP_HL = 16;    // halo either side (>= KP / 2)
constexpr int LDBP_EXT = LDBP_CH + 2 * LDBP_HL;
constexpr int LDBP_RB = 4;     // consecutive samples per work unit
constexpr int LDBP_KB = 8;     // taps per register window

__host__ __device__ constexpr int ldbp_bwd_threads(int modes) {
    return ((modes * LDBP_EXT / LDBP_RB + 63) / 64) * 64;
}

template <int K, int KP, int M>
__global__ __launch_bounds__(ldbp_bwd_threads(M)) void k_ldbp_bwd_step(LdbpArgs a, int j, const cx<float>* gy_in,
                                                                        cx<float>* gy_out) {
    constexpr int NT = ldbp_bwd_threads(M);
    constexpr int R = LDBP_RB, KC = LDBP_KB, P = KP / 2, c = K / 2;
    constexpr int CH = LDBP_CH, HL = LDBP_HL, EXT = LDBP_EXT;
    constexpr int KR = (K + KC - 1) / KC * KC;  // taps rounded up to whole windows
    constexpr int WIN = EXT + KR + R;           // G_y window of the halo-wide input gradient
    constexpr int YW = EXT + 2 * P;             // y window of the intensities on chunk and halo
    constexpr int XW = CH + KR + R;             // x window of the tap gradients
    static_assert(P <= HL, "the halo covers the intensity filter");
    static_assert(EXT % R == 0 && CH % R == 0, "whole units");
    __shared__ cx<float> gyw[M][WIN];  // G_y of step j + 1 at lo - HL - c + w
    __shared__ cx<float> gz[M][EXT];   // G_z of step j at lo - HL + e
    __shared__ cx<float> yw[M][YW];    // y of step j at lo - HL - P + v
    __shared__ cx<float> xw[M][XW];    // x of step j at lo + c - (KR - 1) - ... (see below)
    __shared__ float2 qphi[EXT];       // q and phi at lo - HL + e
    __shared__ float pw[YW];           // intensity summed over the modes, as yw
    constexpr int TP = NT / KR;        // slices of the chunk per FIR tap
    constexpr int MP = NT / KP;        // slices per intensity tap
    __shared__ double hsum[TP][KR][2];  // per-slice FIR tap gradients
    __shared__ double psum[MP][KP];     // per-slice intensity-tap gradients
    const int N = a.N, mask = N - 1;
    const int nch = N / CH;
    const int b = blockIdx.x / nch, lo = (blockIdx.x % nch) * CH;
    const long long wb = (long long)b * M * N;
    const cx<float>* x = a.xs + ((long long)b * a.J + j) * M * N;  // [M][N] FIR input of step j
    const cx<float>* y = a.ys + ((long long)b * a.J + j) * M * N;  // [M][N] FIR output of step j
    const float* gk = a.kerr + j * KP;
    const float* hj = reinterpret_cast<const float*>(a.fir + (long long)j * K);
    // x window: xw[v] = x(lo + c - KR + 1 - ... ) laid out so that unit u0's
    // tap k reads xw[u0 + r - k + KR - 1]: xw[v] = x(lo + c - (KR - 1) + v)
    const int x0 = lo + c - (KR - 1);

    // stage the windows: every load issued before the first LDS write, so
    // that their latencies overlap
    constexpr int IY = (M * YW + NT - 1) / NT, IX = (M * XW + NT - 1) / NT;
    constexpr int IG = (M * WIN + NT - 1) / NT, IZ = (M * EXT + NT - 1) / NT;
    const bool last = j == a.J - 1;
    cx<float> sy[IY], sx[IX], sg[IG > 2 * IZ ? IG : 2 * IZ];
#pragma unroll
    for (int k = 0; k < IY; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * YW) sy[k] = y[(i / YW) * N + ((lo - HL - P + i % YW) & mask)];
    }
#pragma unroll
    for (int k = 0; k < IX; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * XW) sx[k] = x[(i / XW) * N + ((x0 + i % XW) & mask)];
    }
    if (last) {
#pragma unroll
        for (int k = 0; k < IZ; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < M * EXT) {
                const long long t = wb + (long long)(i / EXT) * N + ((lo - HL + i % EXT) & mask);
                sg[2 * k] = a.out[t];
                sg[2 * k + 1] = a.target[t];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < IG; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < M * WIN) sg[k] = gy_in[wb + (long long)(i / WIN) * N + ((lo - HL - c + i % WIN) & mask)];
        }
    }
#pragma unroll
    for (int k = 0; k < IY; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * YW) yw[i / YW][i % YW] = sy[k];
    }
#pragma unroll
    for (int k = 0; k < IX; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * XW) xw[i / XW][i % XW] = sx[k];
    }
    if (last) {
        const float scale = 1.0f / ((float)N * (float)M * (float)a.B);
#pragma unroll
        for (int k = 0; k < IZ; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < M * EXT) gz[i / EXT][i % EXT] = (sg[2 * k] - sg[2 * k + 1]) * scale;
        }
    } else {
#pragma unroll
        for (int k = 0; k < IG; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < M * WIN) gyw[i / WIN][i % WIN] = sg[k];
        }
    }
    __syncthreads();

    for (int v = threadIdx.x; v < YW; v += NT) {
        float pv = norm2(yw[0][v]);
#pragma unroll
        for (int md = 1; md < M; ++md) pv += norm2(yw[md][v]);
        pw[v] = pv;
    }

    // 1. G_z of step j as the input gradient of step j + 1 on chunk and halo:
    // G_x(s) = sum_k conj(h_k) G_y(s + k - c), s = lo - HL + e, from
    // gyw[e + k]
    if (j < a.J - 1) {
        const float* hn = reinterpret_cast<const float*>(a.fir + (long long)(j + 1) * K);
        const int u = threadIdx.x;
        if (u < M * (EXT / R)) {
            const int md = u / (EXT / R), e0 = (u % (EXT / R)) * R;
            cx<float> gx[R];
#pragma unroll
            for (int r = 0; r < R; ++r) gx[r] = mk<float>(0, 0);
#pragma unroll 1
            for (int k0 = 0; k0 < K; k0 += KC) {
                cx<float> w[R + KC - 1];  // G_y at window positions e0 + k0 + v
#pragma unroll
                for (int v = 0; v < R + KC - 1; ++v) w[v] = gyw[md][e0 + k0 + v];
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const bool live = k0 + q < K;
                    const float hr = live ? sload(hn + 2 * (k0 + q)) : 0.0f;
                    const float hi = live ? sload(hn + 2 * (k0 + q) + 1) : 0.0f;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const cx<float> gv = w[r + q];  // conj(h) * G_y
                        gx[r].x = fmaf(hr, gv.x, fmaf(hi, gv.y, gx[r].x));
                        gx[r].y = fmaf(hr, gv.y, fmaf(-hi, gv.x, gx[r].y));
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) gz[md][e0 + r] = gx[r];
        }
    }
    __syncthreads();

    // 2. q and phi on chunk and halo
    for (int e = threadIdx.x; e < EXT; e += NT) {
        float phi = 0;  // sum_m g_m p(t - m + P), p(t - m + P) = pw[e + 2 P - m]
#pragma unroll
        for (int m = 0; m < KP; ++m) phi = fmaf(gk[m], pw[e + 2 * P - m], phi);
        float s, co;
        sincos_fast(phi, &s, &co);
        float q = 0;  // -sum_modes Im(conj(gz) z)
#pragma unroll
        for (int md = 0; md < M; ++md) {
            const cx<float> zv = yw[md][e + P] * mk<float>(co, s);
            const cx<float> gzv = gz[md][e];
            q -= gzv.x * zv.y - gzv.y * zv.x;
        }
        qphi[e] = make_float2(q, phi);
    }
    __syncthreads();

    // 3. G_y = G_z e^{-i phi} + (dL/dp) y on the chunk, kept in gz and
    // written out for the next launch
    for (int i = threadIdx.x; i < CH; i += NT) {
        const int e = HL + i;
        float dp = 0;
#pragma unroll
        for (int m = 0; m < KP; ++m) dp = fmaf(gk[m], 2.0f * qphi[e + m - P].x, dp);
        float s, co;
        sincos_fast(qphi[e].y, &s, &co);
#pragma unroll
        for (int md = 0; md < M; ++md) {
            const cx<float> gv = gz[md][e] * mk<float>(co, -s) + yw[md][e + P] * dp;
            gz[md][e] = gv;
            gy_out[wb + (long long)md * N + lo + i] = gv;
        }
    }
    __syncthreads();

    // 4. parameter gradients over the chunk, each as whole sums of one
    // parameter over one slice of the chunk (no cross-lane reductions):
    // FIR taps G_h_k = sum_modes sum_t G_y(t) conj(x(t - k + c)), with
    // x(t - k + c) = xw[t - lo - k + KR - 1], and intensity taps
    // dL/dg_m = sum_t 2 q(t) p(t - m + P)
    {
        const int k = threadIdx.x % KR, sl = threadIdx.x / KR;
        if (sl < TP) {
            constexpr int SPAN = (CH + TP - 1) / TP;
            const int i0 = sl * SPAN, i1 = i0 + SPAN < CH ? i0 + SPAN : CH;
            float dre = 0, dim = 0;
#pragma unroll
            for (int md = 0; md < M; ++md) {
                for (int i = i0; i < i1; ++i) {
                    const cx<float> gv = gz[md][HL + i];
                    const cx<float> xv = xw[md][i - k + KR - 1];
                    dre = fmaf(gv.x, xv.x, fmaf(gv.y, xv.y, dre));
                    dim = fmaf(gv.y, xv.x, fmaf(-gv.x, xv.y, dim));
                }
            }
            hsum[sl][k][0] = dre;
            hsum[sl][k][1] = dim;
        }
        const int m = threadIdx.x % KP, sp = threadIdx.x / KP;
        if (sp < MP) {
            constexpr int SPAN = (CH + MP - 1) / MP;
            const int i0 = sp * SPAN, i1 = i0 + SPAN < CH ? i0 + SPAN : CH;
            double d = 0;
            for (int i = i0; i < i1; ++i) d += 2.0 * (double)qphi[HL + i].x * (double)pw[HL + i + 2 * P - m];
            psum[sp][m] = d;
        }
    }
    // this workgroup's sums of the step's parameter gradients, slices added
    // in order, in its own slot (k_ldbp_grad_reduce adds the slots in a
    // fixed order)
    __syncthreads();
    double* slot = a.part + ((long long)j * gridDim.x + blockIdx.x) * (2 * K + KP);
    for (int i = threadIdx.x; i < 2 * K + KP; i += NT) {
        double st = 0;
        if (i < 2 * K) {
            for (int sl = 0; sl < TP; ++sl) st += hsum[sl][i >> 1][i & 1];
        } else {
            for (int sp = 0; sp < MP; ++sp) st += psum[sp][i - 2 * K];
        }
        slot[i] = st;
    }
}

// Training forward pass split over the whole device: one launch per step j,
// one workgroup per LDBP_CH-sample chunk of one waveform (all modes).  The
// step reads its input x_j (the caller's input for j = 0, else xs[j]), forms
// the FIR output y on the chunk and a LDBP_HL-sample halo either side, the
// phase from the filtered intensity and z = y e^{i phi} on the chunk, and
// stores y (ys[j]) and z (xs[j + 1], or out after the last step, with the
// chunk's share of the loss in its own slot of lslot).
template <int K, int KP, int M>
__global__ __launch_bounds__(ldbp_bwd_threads(M)) void k_ldbp_fwd_step(LdbpArgs a, int j, double* lslot) {
    constexpr int NT = ldbp_bwd_threads(M);
    constexpr int R = LDBP_RB, KC = LDBP_KB, P = KP / 2, c = K / 2;
    constexpr int CH = LDBP_CH, HL = LDBP_HL, EXT = LDBP_EXT;
    constexpr int KR = (K + KC - 1) / KC * KC;
    constexpr int XW = EXT + KR + R;  // x window: xw[v] = x(lo - HL + c - (KR - 1) + v)
    static_assert(P <= HL, "the halo covers the intensity filter");
    __shared__ cx<float> xw[M][XW];
    __shared__ cx<float> yv[M][EXT];  // y at lo - HL + e
    __shared__ float pw[EXT];         // intensity summed over the modes
    __shared__ double red[NT / 64];
    const int N = a.N, mask = N - 1;
    const int nch = N / CH;
    const int b = blockIdx.x / nch, lo = (blockIdx.x % nch) * CH;
    const long long wb = (long long)b * M * N;
    const cx<float>* x = j == 0 ? a.in + wb : a.xs + ((long long)b * a.J + j) * M * N;
    const int x0 = lo - HL + c - (KR - 1);
    constexpr int IX = (M * XW + NT - 1) / NT;
    cx<float> sx[IX];
#pragma unroll
    for (int k = 0; k < IX; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * XW) sx[k] = x[(i / XW) * N + ((x0 + i % XW) & mask)];
    }
#pragma unroll
    for (int k = 0; k < IX; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < M * XW) xw[i / XW][i % XW] = sx[k];
    }
    if (j == 0 && a.xs) {  // the backward pass reads every step's input
        cx<float>* xs0 = a.xs + (long long)b * a.J * M * N;
        for (int i = threadIdx.x; i < M * CH; i += NT) {
            const int md = i / CH, t = lo + i % CH;
            xs0[md * N + t] = x[md * N + t];
        }
    }
    __syncthreads();
    // FIR on chunk and halo, units of R outputs: y(t) = sum_k h_k x(t - k + c),
    // x(t - k + c) = xw[e - k + KR - 1] for t = lo - HL + e
    const float* hj = reinterpret_cast<const float*>(a.fir + (long long)j * K);
    const int u = threadIdx.x;
    if (u < M * (EXT / R)) {
        const int md = u / (EXT / R), e0 = (u % (EXT / R)) * R;
        cx<float> yr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) yr[r] = mk<float>(0, 0);
#pragma unroll 1
        for (int k0 = 0; k0 < K; k0 += KC) {
            cx<float> w[R + KC - 1];
            const int base = e0 - k0 + KR - KC;
#pragma unroll
            for (int v = 0; v < R + KC - 1; ++v) w[v] = xw[md][base + v];
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                const bool live = k0 + q < K;
                const float hr = live ? sload(hj + 2 * (k0 + q)) : 0.0f;
                const float hi = live ? sload(hj + 2 * (k0 + q) + 1) : 0.0f;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const cx<float> xv = w[r - q + KC - 1];
                    yr[r].x = fmaf(hr, xv.x, fmaf(-hi, xv.y, yr[r].x));
                    yr[r].y = fmaf(hr, xv.y, fmaf(hi, xv.x, yr[r].y));
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) yv[md][e0 + r] = yr[r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EXT; e += NT) {
        float pv = norm2(yv[0][e]);
#pragma unroll
        for (int md = 1; md < M; ++md) pv += norm2(yv[md][e]);
        pw[e] = pv;
    }
    __syncthreads();
    // phase and output on the chunk
    const float* gk = a.kerr + j * KP;
    const bool last = j == a.J - 1;
    cx<float>* ys = a.ys + ((long long)b * a.J + j) * M * N;
    cx<float>* z = last ? a.out + wb : a.xs + ((long long)b * a.J + j + 1) * M * N;
    double part = 0;
    for (int i = threadIdx.x; i < CH; i += NT) {
        const int e = HL + i, t = lo + i;
        float phi = 0;
#pragma unroll
        for (int m = 0; m < KP; ++m) phi = fmaf(gk[m], pw[e - m + P], phi);
        float s, co;
        sincos_fast(phi, &s, &co);
#pragma unroll
        for (int md = 0; md < M; ++md) {
            const cx<float> yy = yv[md][e];
            const cx<float> zz = yy * mk<float>(co, s);
            ys[md * N + t] = yy;
            z[md * N + t] = zz;
            if (last && a.target) part += (double)norm2(zz - a.target[wb + md * N + t]);
        }
    }
    if (last && a.target) {
        part = ldbp_wave_sum(part);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0;
            for (int w = 0; w < NT / 64; ++w) sum += red[w];
            lslot[blockIdx.x] = sum / ((double)N * M * a.B);
        }
    }
}

// *loss += the chunks' loss shares, added in a fixed order
__global__ __launch_bounds__(256) void k_ldbp_loss_reduce(const double* lslot, int groups, double* loss) {
    __shared__ double red[4];
    double s = 0;
    for (int g = threadIdx.x; g < groups; g += 256) s += lslot[g];
    s = ldbp_wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *loss += (red[0] + red[1]) + (red[2] + red[3]);
}

// grad += the per-workgroup sums of every step, one workgroup per (step,
// parameter), the groups' slots added in a fixed order: the gradient does
// not depend on the order in which workgroups finish
__global__ __launch_bounds__(256) void k_ldbp_grad_reduce(const double* part, int groups, int J, int K, int KP,
                                                          double* grad) {
    __shared__ double red[4];
    const int np = 2 * K + KP;
    const int j = blockIdx.x / np, i = blockIdx.x % np;
    const double* p = part + (long long)j * groups * np + i;
    double s = 0;
    for (int g = threadIdx.x; g < groups; g += 256) s += p[(long long)g * np];
    s = ldbp_wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = (red[0] + red[1]) + (red[2] + red[3]);
        double* dst = i < 2 * K ? grad + (long long)j * K * 2 + i : grad + (long long)J * K * 2 + (long long)j * KP + (i - 2 * K);
        *dst += t;
    }
}

// ------------------------------------------------------------------
// Inference on the matrix cores.  The FIR of one step is a block-Toeplitz
// product: 16 consecutive outputs y[16 b + m] of block b take the window
// x[w0(b) + q], w0(b) = 16 b + c - (K - 1) - DELTA, through
//     T[m][q] = h[m - q + K - 1 + DELTA]   (zero outside 0..K-1).
// In real form, rows 2m + e (e: re/im of the output) and columns 2q + f (f:
// re/im of the input) hold [[h_r, -h_i], [h_i, h_r]]; 32 blocks form the
// 32 columns of one v_mfma_f32_32x32x16_f16 tile (512 outputs), and the
// window is covered by NKS k-steps of 16.  Operands are FP16 with FP32
// accumulation; the field is held in LDS as FP16 between steps (rounding
// adds ~2.8e-4 relative noise per step, about -52 dB over 80 steps; BF16's
// 8-bit mantissa would add ~-34 dB, enough to cost Q at 24 dB), the
// nonlinear phase is computed in FP32 from the FP16 FIR output, and the last
// step writes FP32.
// Lane layout of the 32x32x16 MFMA (checked on the device): lane l holds A
// row l % 32 and B column l % 32 at k = 8 (l / 32) + u, u < 8, and
// accumulator v is D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
// Neighbouring lanes own neighbouring blocks, so the field and the
// intensities are stored block-transposed in LDS (sample 16 b + m at row m,
// column b): lanes then touch consecutive words, in the matrix phase and in
// the nonlinear phase, where one thread owns one block.
// ------------------------------------------------------------------
typedef _Float16 ldbp_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 ldbp_h2 __attribute__((ext_vector_type(2)));  // two FP16 lanes of one packed instruction
typedef float ldbp_f16 __attribute__((ext_vector_type(16)));

struct cx_h {
    _Float16 x, y;
};
// two consecutive samples
struct __align__(8) cx_h2 {
    cx_h lo, hi;
};
// four consecutive samples (rows 4 rq .. 4 rq + 3 of one block); their bytes
// are the FP16 B operand of one matrix-core product
struct __align__(16) cx_h4 {
    cx_h s[4];
};

constexpr int LI_NT = 1024;  // 16 waves; each wave owns up to 2 tiles of 512 outputs

template <int K>
struct InferShape {
    static constexpr int c = K / 2;
    static constexpr int DELTA = ((c - (K - 1)) % 4 + 4) % 4;
    static constexpr int NKS = (16 + K - 1 + DELTA + 7) / 8;
};

// A operands of every step: afrag[(j NKS + s) 64 + lane] (8 FP16 each)
template <int K>
__global__ void k_ldbp_afrag(const cx<float>* __restrict__ fir, int J, ldbp_h8* __restrict__ afrag) {
    using IS = InferShape<K>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= J * IS::NKS * 64) return;
    const int lane = i & 63, s = (i >> 6) % IS::NKS, j = (i >> 6) / IS::NKS;
    const int hl = lane >> 5, col = lane & 31, m_row = col >> 1, e_row = col & 1;
    ldbp_h8 A;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int q = 8 * s + 4 * hl + (u >> 1);
        const int k = m_row - q + K - 1 + IS::DELTA;
        float v = 0.0f;
        if (k >= 0 && k < K) {
            const cx<float> hk = fir[j * K + k];
            v = e_row == 0 ? ((u & 1) == 0 ? hk.x : -hk.y) : ((u & 1) == 0 ? hk.y : hk.x);
        }
        A[u] = (_Float16)v;
    }
    afrag[i] = A;
}

// M modes: LDS holds the field (FP16) and each mode's intensities, both as
// [M][4][nb] quads of consecutive samples: quad (rq, blk) holds samples
// 16 blk + 4 rq .. 16 blk + 4 rq + 3, so one 16-byte read is one operand of
// the FIR product and neighbouring lanes (blocks) stay in neighbouring banks.
// The M N / 512 tiles (at most 32: M N <= 16384) are dealt over the 16 waves,
// and the nonlinear phase sums the modes.
template <int K, int KP, int M>
__global__ __launch_bounds__(LI_NT) void k_ldbp_infer(LdbpArgs a, const ldbp_h8* __restrict__ afrag) {
    using IS = InferShape<K>;
    constexpr int c = IS::c, DELTA = IS::DELTA, NKS = IS::NKS, P = KP / 2;
    constexpr int C0 = c - (K - 1) - DELTA;  // operand window start relative to 16 blk, half 0
    static_assert(C0 % 4 == 0, "the four samples of an operand are one quad");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = a.N, nb = N >> 4, lnb = __builtin_ctz(N >> 4), NQ = N >> 2;
    cx_h4* xs = reinterpret_cast<cx_h4*>(smem_raw);                                   // [M][4][nb]
    float4* pb = reinterpret_cast<float4*>(smem_raw + (size_t)M * N * sizeof(cx_h));  // [M][4][nb]
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int hl = lane >> 5, col = lane & 31;
    const int mode_tiles = N / 512, n_tiles = M * mode_tiles;
    const int b = blockIdx.x;
    const cx<float>* in = a.in + (long long)b * M * N;
#pragma unroll
    for (int md = 0; md < M; ++md)
        for (int q = threadIdx.x; q < NQ; q += LI_NT) {
            const int t = ((q & (nb - 1)) << 4) + ((q >> lnb) << 2);  // first sample of quad q
            cx_h4 h;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const cx<float> v = in[(long long)md * N + t + e];
                h.s[e] = cx_h{(_Float16)v.x, (_Float16)v.y};
            }
            xs[md * NQ + q] = h;
        }
    __syncthreads();
    // the matrix-core operands of step j, loaded during step j - 1's phase
    ldbp_h8 A[NKS];
#pragma unroll
    for (int s = 0; s < NKS; ++s) A[s] = afrag[s * 64 + lane];
    for (int j = 0; j < a.J; ++j) {
        // FIR on the matrix cores
        ldbp_f16 acc[2];
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            acc[it] = ldbp_f16{};
            const int tile = wid + 16 * it;
            if (tile < n_tiles) {
                // lane coordinates made opaque per step: the operand
                // addresses are formed here rather than held across the
                // steps, where they would be spilled
                int colj = col, hlj = hl;
                __asm__ volatile("" : "+v"(colj), "+v"(hlj));
                const cx_h4* xm = xs + (tile / mode_tiles) * NQ;
                const int blk = (tile % mode_tiles) * 32 + colj;
#pragma unroll
                for (int s = 0; s < NKS; ++s) {
                    // samples 16 blk + o .. + 3, o = C0 + 4 hl + 8 s: quad
                    // (o & 15) / 4 of block blk + (o >> 4)
                    const int oa = C0 + 8 * s, ob = oa + 4;
                    const int rq = hlj ? ((ob & 15) >> 2) : ((oa & 15) >> 2);
                    const int bo = hlj ? (ob >> 4) : (oa >> 4);
                    const ldbp_h8 bf = __builtin_bit_cast(ldbp_h8, xm[(rq << lnb) + ((blk + bo) & (nb - 1))]);
                    acc[it] = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[s], bf, acc[it], 0, 0, 0);
                }
            }
        }
        // every read of the field is done before it is overwritten
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int tile = wid + 16 * it;
            if (tile < n_tiles) {
                const int mo = (tile / mode_tiles) * NQ;
                const int blk = (tile % mode_tiles) * 32 + col;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // rows 4 i + 2 hl and 4 i + 2 hl + 1: half hl of quad i
                    const float r0 = acc[it][4 * i], i0 = acc[it][4 * i + 1];
                    const float r1 = acc[it][4 * i + 2], i1 = acc[it][4 * i + 3];
                    const int q = mo + (i << lnb) + blk;
                    reinterpret_cast<cx_h2*>(xs + q)[hl] =
                        cx_h2{cx_h{(_Float16)r0, (_Float16)i0}, cx_h{(_Float16)r1, (_Float16)i1}};
                    reinterpret_cast<float2*>(pb + q)[hl] = make_float2(r0 * r0 + i0 * i0, r1 * r1 + i1 * i1);
                }
            }
        }
        __syncthreads();
        // the next step's operands, in flight during this step's phase
        if (j + 1 < a.J) {
#pragma unroll
            for (int s = 0; s < NKS; ++s) A[s] = afrag[((j + 1) * NKS + s) * 64 + lane];
        }
        // nonlinear phase, one thread per block of 16 samples
        const bool last = j == a.J - 1;
        float g[KP];
        ldbp_h2 g2[KP];
#pragma unroll
        for (int mm = 0; mm < KP; ++mm) {
            g[mm] = sload(a.kerr + j * KP + mm);
            // the packed pair is wave-uniform: held in a scalar register
            const ldbp_h2 gv = ldbp_h2{(_Float16)g[mm], (_Float16)g[mm]};
            g2[mm] = __builtin_bit_cast(
                ldbp_h2, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, gv)));
        }
        for (int blk = threadIdx.x; blk < nb; blk += LI_NT) {
            // intensities of quad rq of block bl, summed over the modes
            auto inten4 = [&](int rq, int bl) {
                float4 v = pb[(rq << lnb) + bl];
#pragma unroll
                for (int md = 1; md < M; ++md) {
                    const float4 u = pb[md * NQ + (rq << lnb) + bl];
                    v.x += u.x;
                    v.y += u.y;
                    v.z += u.z;
                    v.w += u.w;
                }
                return v;
            };
            // pe[u] = p(16 blk - P + u), summed over the modes
            auto gather = [&](float (&pe)[16 + 2 * P]) {
#pragma unroll
                for (int rq = 0; rq < 4; ++rq) {
                    const float4 v = inten4(rq, blk);
                    pe[P + 4 * rq] = v.x;
                    pe[P + 4 * rq + 1] = v.y;
                    pe[P + 4 * rq + 2] = v.z;
                    pe[P + 4 * rq + 3] = v.w;
                }
                if constexpr (P > 0) {
                    const int bp = (blk + nb - 1) & (nb - 1), bn = (blk + 1) & (nb - 1);
                    // rows 16 - P .. 15 of the block before, rows 0 .. P - 1 of the one after
                    float lo[16], hi[16];
#pragma unroll
                    for (int rq = 0; rq < 4; ++rq) {
                        if (4 * rq + 3 >= 16 - P) {
                            const float4 v = inten4(rq, bp);
                            lo[4 * rq] = v.x;
                            lo[4 * rq + 1] = v.y;
                            lo[4 * rq + 2] = v.z;
                            lo[4 * rq + 3] = v.w;
                        }
                        if (4 * rq < P) {
                            const float4 v = inten4(rq, bn);
                            hi[4 * rq] = v.x;
                            hi[4 * rq + 1] = v.y;
                            hi[4 * rq + 2] = v.z;
                            hi[4 * rq + 3] = v.w;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        pe[i] = lo[16 - P + i];
                        pe[P + 16 + i] = hi[i];
                    }
                }
            };
            // Steps before the last whose 16 phases all lie within 1/2 rad
            // (wave-uniform; the learned models' per-step phases) run on
            // packed FP16, two samples per instruction: the intensity filter
            // over pairs of samples, sin and cos - 1 as polynomials (x^5,
            // x^4), and the rotation as z + z (cos - 1 + i sin), so the
            // FP16 rounding of the small correction stays below 1e-4 of |z|
            // and the field is rounded once, as the FP32 path rounds it.
            bool fast = !last;
            ldbp_h2 cm2[8], sn2[8];
            if (fast) {
                ldbp_h2 p2[8 + P];  // (p_2k, p_2k+1) of pe
                ldbp_h2 po[7 + P];  // (p_2k+1, p_2k+2)
                {
                    float pe[16 + 2 * P];
                    gather(pe);
#pragma unroll
                    for (int k = 0; k < 8 + P; ++k) p2[k] = ldbp_h2{(_Float16)pe[2 * k], (_Float16)pe[2 * k + 1]};
                }
#pragma unroll
                for (int k = 0; k < 7 + P; ++k) po[k] = __builtin_shufflevector(p2[k], p2[k + 1], 1, 2);
                ldbp_h2 mx2 = {0, 0};
#pragma unroll
                for (int r2 = 0; r2 < 8; ++r2) {
                    ldbp_h2 phi = {0, 0};
#pragma unroll
                    for (int mm = 0; mm < KP; ++mm) {
                        const int base = 2 * r2 - mm + 2 * P;  // samples 2 r2, 2 r2 + 1 take p_base, p_base+1
                        const ldbp_h2 pv = (base & 1) ? po[(base - 1) >> 1] : p2[base >> 1];
                        phi = g2[mm] * pv + phi;
                    }
                    sn2[r2] = phi;
                    mx2 = __builtin_elementwise_max(mx2, __builtin_elementwise_abs(phi));
                }
                fast = __all((float)__builtin_elementwise_max(mx2.x, mx2.y) <= 0.5f);
            }
            if (fast) {
#pragma unroll
                for (int r2 = 0; r2 < 8; ++r2) {
                    const ldbp_h2 x = sn2[r2], x2 = x * x;
                    const ldbp_h2 t = x2 * ldbp_h2{(_Float16)(1.0f / 120), (_Float16)(1.0f / 120)} +
                                      ldbp_h2{(_Float16)(-1.0f / 6), (_Float16)(-1.0f / 6)};
                    sn2[r2] = (x * x2) * t + x;
                    cm2[r2] = x2 * (x2 * ldbp_h2{(_Float16)(1.0f / 24), (_Float16)(1.0f / 24)} +
                                    ldbp_h2{(_Float16)-0.5f, (_Float16)-0.5f});
                }
#pragma unroll
                for (int rq = 0; rq < 4; ++rq) {
#pragma unroll
                    for (int md = 0; md < M; ++md) {
                        const int q = md * NQ + (rq << lnb) + blk;
                        ldbp_h8 zq = __builtin_bit_cast(ldbp_h8, xs[q]);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int r = 4 * rq + e;
                            const _Float16 cmr = (r & 1) ? cm2[r >> 1].y : cm2[r >> 1].x;
                            const _Float16 snr = (r & 1) ? sn2[r >> 1].y : sn2[r >> 1].x;
                            const ldbp_h2 z = ldbp_h2{zq[2 * e], zq[2 * e + 1]};
                            ldbp_h2 d = z * ldbp_h2{cmr, cmr};
                            d = z.yx * ldbp_h2{-snr, snr} + d;
                            const ldbp_h2 zn = z + d;
                            zq[2 * e] = zn.x;
                            zq[2 * e + 1] = zn.y;
                        }
                        xs[q] = __builtin_bit_cast(cx_h4, zq);
                    }
                }
                continue;
            }
            float pe[16 + 2 * P];
            gather(pe);
            // otherwise eight phases at a time in FP32, then one wave-uniform
            // choice of the rotation: every phase within 1/8 rad takes a
            // polynomial (sin to x^5, cos to x^4: below 1e-8 off), within
            // 1000 rad the hardware sine and cosine (v_sin_f32 / v_cos_f32,
            // ~1e-6), else the reduced series
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                float sn[8], cs[8], mx = 0.0f;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int r = 8 * half + u;
                    float phi = 0.0f;
#pragma unroll
                    for (int mm = 0; mm < KP; ++mm) phi = fmaf(g[mm], pe[r - mm + 2 * P], phi);
                    sn[u] = phi;
                    mx = fmaxf(mx, fabsf(phi));
                }
                if (__all(mx <= 0.125f)) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float x = sn[u], x2 = x * x;
                        sn[u] = fmaf(x * x2, fmaf(x2, 0.008333333333333333f, -0.16666666666666666f), x);
                        cs[u] = fmaf(x2, fmaf(x2, 0.041666666666666664f, -0.5f), 1.0f);
                    }
                } else if (__all(mx <= 1000.0f)) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float x = sn[u];
                        sn[u] = __sinf(x);
                        cs[u] = __cosf(x);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float x = sn[u];
                        sincos_fast(x, &sn[u], &cs[u]);
                    }
                }
#pragma unroll
                for (int rh = 0; rh < 2; ++rh) {
                    const int rq = 2 * half + rh;
#pragma unroll
                    for (int md = 0; md < M; ++md) {
                        const int q = md * NQ + (rq << lnb) + blk;
                        cx_h4 yh = xs[q];
                        cx<float> z[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            z[e] = mk<float>((float)yh.s[e].x, (float)yh.s[e].y) *
                                   mk<float>(cs[4 * rh + e], sn[4 * rh + e]);
                        if (last) {
                            cx<float>* o = a.out + ((long long)b * M + md) * N + 16 * blk + 4 * rq;
#pragma unroll
                            for (int e = 0; e < 4; ++e) o[e] = z[e];
                        } else {
#pragma unroll
                            for (int e = 0; e < 4; ++e) yh.s[e] = cx_h{(_Float16)z[e].x, (_Float16)z[e].y};
                            xs[q] = yh;
                        }
                    }
                }
            }
        }
        // the new field is visible before the next step's reads
        if (!last) __syncthreads();
    }
}

// Adam (Kingma & Ba, 2015) on float parameters with double gradients (the
// learned DBP's FIR and intensity taps, the triplet MLP's weights).
__global__ void k_adam(float* __restrict__ p, const double* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float beta1, float beta2, float eps, int t,
                       const float* __restrict__ lr_scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // learned DBP: the taps carry the Wirtinger gradient dL/d(conj h) =
    // (dL/dRe + i dL/dIm) / 2; Adam normalises every coordinate, so the
    // factor 2 against the real intensity-tap gradients does not change the
    // update
    float gi = (float)g[i];
    float mi = beta1 * m[i] + (1 - beta1) * gi;
    float vi = beta2 * v[i] + (1 - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / (1 - powf(beta1, (float)t));
    float vh = vi / (1 - powf(beta2, (float)t));
    float s = lr_scale ? lr_scale[i] : 1.0f;
    p[i] -= lr * s * mh / (sqrtf(vh) + eps);
}

}  // namespace focss_hip
