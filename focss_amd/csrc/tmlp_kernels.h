// Triplet MLP post-equaliser (no counterpart in the reference; it widens the
// perturbation-triplet equaliser of src/focss/equalizer/sino.cpp into a
// two-layer network).  For symbol s of mode md of one waveform (symbols
// periodic, indices mod S):
//     T_t(s) = x_md(s + m) sum_p x_p(s + n) conj(x_p(s + m + n)),
//              t = (m + 16) 33 + (n + 16),  |m|, |n| <= 16     (1089 triplets,
//              the term of sino.cpp:154-168 over the full square)
//     f = (Re T_0, Im T_0, Re T_1, ...), padded with zeros to 2240 features
//     pre = W1 f + b1   (W1 [128][2240]),   h = max(pre, 0)
//     y = x + (o_0 + i o_1),   o = W2 h + b2   (W2 [2][128])
// Loss = (1/(B M S)) sum |y - target|^2.  Gradients:
//     dL/do = 2 (y - target) / (B M S),   dpre = (pre > 0) W2^T dL/do,
//     dW1 = dpre f^T,  db1 = sum dpre,  dW2 = dL/do h^T,  db2 = sum dL/do.
// The two products with 2240-long rows (W1 f and dpre f^T) run on the matrix
// cores: BF16 operands (features, W1, dpre), FP32 accumulation.  The features
// are never stored: each kernel forms them from the symbols it holds in LDS.
#pragma once
#include "cx.h"

namespace focss_hip {

constexpr int TM_R = 16;
constexpr int TM_W = 2 * TM_R + 1;          // 33
constexpr int TM_T = TM_W * TM_W;           // 1089 triplets
constexpr int TM_TP = 1120;                 // triplet slots (35 blocks of 32)
constexpr int TM_KF = 2 * TM_TP;            // 2240 features, the W1 row stride
constexpr int TM_H = 128;                   // hidden units
constexpr int TM_NP = TM_H * TM_KF + TM_H + 2 * TM_H + 2;  // parameters
constexpr int TM_SMALL = TM_H + 2 * TM_H + 2 + 1;          // b1, W2, b2 gradients and the loss
constexpr int TM_ROWS = 128;                // symbols per workgroup block
constexpr int TM_WIN = TM_ROWS + 4 * TM_R;  // window: 2R symbols either side
constexpr int TM_QW = TM_ROWS + 2 * TM_R;   // positions of the pair table (window positions R .. R + QW - 1)
constexpr int TM_NT = 256;                  // 4 waves, 32 symbols each
constexpr int TM_KC = 64;                   // features per staged W1 chunk
constexpr int TM_LDW = TM_KC + 8;           // padded LDS row of the chunk (bf16)
constexpr int TM_SPLITS = 64;               // row splits of the weight-gradient product

typedef __bf16 tm_bf8 __attribute__((ext_vector_type(8)));
typedef float tm_f16 __attribute__((ext_vector_type(16)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned tm_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ tm_bf8 tm_frag(uint4 v) { return __builtin_bit_cast(tm_bf8, v); }

// two floats -> packed bf16 pair (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ unsigned tm_pack2(float lo, float hi) {
    typedef __bf16 tm_bf2 __attribute__((ext_vector_type(2)));
    const tm_bf2 v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(unsigned, v);
}

// Pair table of one window (the window holds every mode of one waveform,
// [M][TM_WIN], position 2R + i is symbol s0 + i):
//     Q_m(c) = sum_p x_p(c) conj(x_p(c + m)),  qt[(m + R) TM_QW + c - R],
// so that the triplet at window position c is T_{m,n}(c) = x_md(c + m) Q_m(c + n).
template <int M>
__device__ __forceinline__ void tm_build_pairs(cx<float>* qt, const cx<float>* xw) {
    for (int i = threadIdx.x; i < TM_W * TM_QW; i += TM_NT) {
        const int m = i / TM_QW - TM_R, c = i % TM_QW + TM_R;
        cx<float> acc = mk<float>(0.0f, 0.0f);
#pragma unroll
        for (int p = 0; p < M; ++p) {
            const cx<float> a = xw[p * TM_WIN + c], b = xw[p * TM_WIN + c + m];
            acc.x = fmaf(a.x, b.x, fmaf(a.y, b.y, acc.x));
            acc.y = fmaf(a.y, b.x, fmaf(-a.x, b.y, acc.y));
        }
        qt[i] = acc;
    }
}

// The rows m_lo and m_lo + 1 of the pair table (the m values of one chunk of
// 32 consecutive triplets): qt2[(m - m_lo) TM_QW + c - R].
template <int M, int WIN = TM_WIN, int QW = TM_QW>
__device__ __forceinline__ void tm_build_pairs2(cx<float>* qt2, const cx<float>* xw, int m_lo) {
    for (int i = threadIdx.x; i < 2 * QW; i += TM_NT) {
        const int m = m_lo + i / QW, c = i % QW + TM_R;
        if (m > TM_R) continue;
        cx<float> acc = mk<float>(0.0f, 0.0f);
#pragma unroll
        for (int p = 0; p < M; ++p) {
            const cx<float> a = xw[p * WIN + c], b = xw[p * WIN + c + m];
            acc.x = fmaf(a.x, b.x, fmaf(a.y, b.y, acc.x));
            acc.y = fmaf(a.y, b.x, fmaf(-a.x, b.y, acc.y));
        }
        qt2[i] = acc;
    }
}

__device__ __forceinline__ cx<float> tm_triplet(const cx<float>* qt, const cx<float>* xm, int c, int m, int n) {
    const cx<float> q = qt[(m + TM_R) * TM_QW + c + n - TM_R];
    const cx<float> u = xm[c + m];
    return mk<float>(u.x * q.x - u.y * q.y, u.x * q.y + u.y * q.x);
}

// every mode of waveform seq / M, symbols s0 - 2R .. s0 + WIN - 2R - 1 (mod S)
template <int M, int WIN = TM_WIN>
__device__ __forceinline__ void tm_load_window(cx<float>* xw, const cx<float>* __restrict__ x, int seq, int s0,
                                               int S) {
    const long long base = (long long)(seq / M) * M * S;
    for (int i = threadIdx.x; i < M * WIN; i += TM_NT) {
        const int p = i / WIN, k = i % WIN;
        const int s = (s0 - 2 * TM_R + k + S) % S;
        xw[i] = x[base + (long long)p * S + s];
    }
}

// W1 (float) -> bf16 copy of the same [128][2240] layout
__global__ void k_tmlp_pack(const float* __restrict__ w1, unsigned short* __restrict__ w1b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w1b[i] = (unsigned short)tm_bf16(w1[i]);
}

struct TmlpArgs {
    const cx<float>* x;       // [B][M][S]
    const cx<float>* target;  // [B][M][S] (training)
    cx<float>* out;           // [B][M][S] (inference)
    const unsigned short* w1b;  // [128][2240] bf16
    const float* b1;          // [128]
    const float* w2;          // [2][128]
    const float* b2;          // [2]
    unsigned short* dpre;     // [128][rows] bf16 (training)
    float* small;             // [TM_SMALL][blocks] per-block partial sums (training)
    float* part;              // [splits][128][2240] weight-gradient partials (training)
    int S, rows, blocks;
    float inv_count;
};

// Forward: one workgroup per 128 consecutive symbols of one (waveform, mode);
// wave w owns symbols 32 w .. 32 w + 31 of the block and all 128 hidden units
// (four 32 x 32 accumulator tiles).  The features are the A operand (row =
// symbol, 8 consecutive features = 4 triplets per lane), W1 chunks staged in
// LDS the B operand.  TRAIN adds the loss and the backward epilogue.
template <int M, bool TRAIN>
__global__ __launch_bounds__(TM_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_tmlp_fwd(TmlpArgs a) {
    __shared__ cx<float> xw[M * TM_WIN];
    __shared__ cx<float> qt2[2 * TM_QW];  // the pair-table rows of the current chunk
    __shared__ __align__(16) unsigned short w1s[TM_H * TM_LDW];
    __shared__ float red[TRAIN ? 4 * TM_SMALL : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, cl = lane & 31, hl = lane >> 5;
    const int blk = blockIdx.x;
    const int row0 = blk * TM_ROWS;
    const int seq = row0 / a.S, s0 = row0 % a.S, md = seq % M;
    tm_load_window<M>(xw, a.x, seq, s0, a.S);
    // W1 chunk c: rows j, features 64 c .. 64 c + 63; thread moves 4 x 16 B
    // (piece q: row (tid + 256 q) / 8, 8 features at 8 ((tid + 256 q) % 8))
    uint4 pw0, pw1, pw2, pw3;
    auto w1_src = [&](int c, int q) {
        const int idx = threadIdx.x + TM_NT * q;
        return reinterpret_cast<const uint4*>(a.w1b + (size_t)(idx >> 3) * TM_KF + TM_KC * c + 8 * (idx & 7));
    };
    auto w1_dst = [&](int q) {
        const int idx = threadIdx.x + TM_NT * q;
        return reinterpret_cast<uint4*>(w1s + (idx >> 3) * TM_LDW + 8 * (idx & 7));
    };
    auto fetch = [&](int c) {
        pw0 = *w1_src(c, 0);
        pw1 = *w1_src(c, 1);
        pw2 = *w1_src(c, 2);
        pw3 = *w1_src(c, 3);
    };
    auto stage = [&]() {
        *w1_dst(0) = pw0;
        *w1_dst(1) = pw1;
        *w1_dst(2) = pw2;
        *w1_dst(3) = pw3;
    };
    // first m of chunk c (32 triplets, 33 per value of m)
    auto chunk_m = [](int c) { return (TM_KC / 2) * c / TM_W - TM_R; };
    fetch(0);
    stage();
    __syncthreads();  // window loaded
    tm_build_pairs2<M>(qt2, xw, chunk_m(0));
    const cx<float>* xm = xw + md * TM_WIN;
    tm_f16 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = tm_f16{};
    const int c_sym = 2 * TM_R + 32 * w + cl;  // window position of this lane's symbol
    constexpr int NCH = TM_KF / TM_KC;          // 35
    __syncthreads();
    for (int c = 0; c < NCH; ++c) {
        if (c + 1 < NCH) fetch(c + 1);
        const int m_lo = chunk_m(c);
        // the lane's triplets of this chunk are t0 + d, d = 8 ks + q < 32 < 33:
        // (m, n) follow from t0's by at most one carry of n into m
        const int t0 = (TM_KC / 2) * c + 4 * hl;
        const int m0 = t0 / TM_W, n0 = t0 - TM_W * m0;  // m = m0 - R (+1 on carry), n likewise
        const int qbase = (m0 - TM_R - m_lo) * TM_QW + c_sym + n0 - 2 * TM_R;
        const int xbase = c_sym + m0 - TM_R;
        // x_md(s + m) takes two values in a chunk: m and m + 1 after a carry
        const cx<float> u_lo = xm[xbase], u_hi = xm[xbase + 1];
#pragma unroll
        for (int ks = 0; ks < TM_KC / 16; ++ks) {
            uint4 av;
            unsigned* ap = reinterpret_cast<unsigned*>(&av);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // branch-free: the padding slots t >= 1089 (last chunk only,
                // m = 17, table row 1) read in-bounds LDS and are zeroed
                const int d = 8 * ks + q;
                const bool carry = n0 + d >= TM_W;
                const cx<float> qv = qt2[qbase + d + (carry ? TM_QW - TM_W : 0)];
                const cx<float> u = carry ? u_hi : u_lo;
                const unsigned v = tm_pack2(u.x * qv.x - u.y * qv.y, u.x * qv.y + u.y * qv.x);
                ap[q] = t0 + d < TM_T ? v : 0u;
            }
            const tm_bf8 A = tm_frag(av);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint4 bv = *reinterpret_cast<const uint4*>(w1s + (32 * t + cl) * TM_LDW + 16 * ks + 8 * hl);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, tm_frag(bv), acc[t], 0, 0, 0);
            }
        }
        if (c + 1 < NCH) {
            __syncthreads();  // every wave is done with chunk c
            stage();
            tm_build_pairs2<M>(qt2, xw, chunk_m(c + 1));
            __syncthreads();
        }
    }
    // epilogue: lane holds pre-activations of hidden j = 32 t + cl for the
    // symbols r(v) = 8 (v / 4) + 4 hl + v % 4 of its wave
    float o0[16], o1[16];
    float b1v[4], w20[4], w21[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        b1v[t] = a.b1[32 * t + cl];
        w20[t] = a.w2[32 * t + cl];
        w21[t] = a.w2[TM_H + 32 * t + cl];
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        float s0v = 0.0f, s1v = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float h = fmaxf(acc[t][v] + b1v[t], 0.0f);
            s0v = fmaf(w20[t], h, s0v);
            s1v = fmaf(w21[t], h, s1v);
        }
        o0[v] = s0v;
        o1[v] = s1v;
    }
    // sum over the 32 hidden columns of each half-wave
#pragma unroll
    for (int off = 1; off < 32; off <<= 1)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            o0[v] += __shfl_xor(o0[v], off, 64);
            o1[v] += __shfl_xor(o1[v], off, 64);
        }
    const float b20 = a.b2[0], b21 = a.b2[1];
    const long long rbase = (long long)row0 + 32 * w;  // global row of the wave's symbol 0
    if constexpr (!TRAIN) {
        if (cl < 16) {
            const int v = cl, r = 8 * (v >> 2) + 4 * hl + (v & 3);
            float y0 = 0.0f, y1 = 0.0f;
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (u == v) {
                    y0 = o0[u];
                    y1 = o1[u];
                }
            const cx<float> xv = xw[md * TM_WIN + 2 * TM_R + 32 * w + r];
            a.out[rbase + r] = mk<float>(xv.x + y0 + b20, xv.y + y1 + b21);
        }
    } else {
        // dL/do for the lane's 16 symbols (every lane of the half holds the same)
        float d0[16], d1[16];
        float loss = 0.0f, g20 = 0.0f, g21 = 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 8 * (v >> 2) + 4 * hl + (v & 3);
            const cx<float> xv = xw[md * TM_WIN + 2 * TM_R + 32 * w + r];
            const cx<float> tg = a.target[rbase + r];
            const float e0 = xv.x + o0[v] + b20 - tg.x, e1 = xv.y + o1[v] + b21 - tg.y;
            loss = fmaf(e0, e0, fmaf(e1, e1, loss));
            d0[v] = 2.0f * e0 * a.inv_count;
            d1[v] = 2.0f * e1 * a.inv_count;
            g20 += d0[v];
            g21 += d1[v];
        }
        float gw0[4], gw1[4], gb1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            gw0[t] = gw1[t] = gb1[t] = 0.0f;
            const int j = 32 * t + cl;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float dpv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int v = 4 * g + e;
                    const float pre = acc[t][v] + b1v[t];
                    const float h = fmaxf(pre, 0.0f);
                    gw0[t] = fmaf(d0[v], h, gw0[t]);
                    gw1[t] = fmaf(d1[v], h, gw1[t]);
                    dpv[e] = pre > 0.0f ? fmaf(w20[t], d0[v], w21[t] * d1[v]) : 0.0f;
                    gb1[t] += dpv[e];
                }
                const unsigned pk[2] = {tm_pack2(dpv[0], dpv[1]), tm_pack2(dpv[2], dpv[3])};
                const long long r = rbase + 8 * g + 4 * hl;
                *reinterpret_cast<uint2*>(a.dpre + (size_t)j * a.rows + r) = make_uint2(pk[0], pk[1]);
            }
            // the two halves cover different symbols of the same hidden unit
            gw0[t] += __shfl_xor(gw0[t], 32, 64);
            gw1[t] += __shfl_xor(gw1[t], 32, 64);
            gb1[t] += __shfl_xor(gb1[t], 32, 64);
        }
        loss += __shfl_xor(loss, 32, 64);
        g20 += __shfl_xor(g20, 32, 64);
        g21 += __shfl_xor(g21, 32, 64);
        float* rw = red + w * TM_SMALL;
        if (hl == 0) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                rw[32 * t + cl] = gb1[t];
                rw[TM_H + 32 * t + cl] = gw0[t];
                rw[2 * TM_H + 32 * t + cl] = gw1[t];
            }
            if (cl == 0) {
                rw[3 * TM_H] = g20;
                rw[3 * TM_H + 1] = g21;
                rw[3 * TM_H + 2] = loss * a.inv_count;
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < TM_SMALL; q += TM_NT)
            a.small[(size_t)q * a.blocks + blk] = red[q] + red[TM_SMALL + q] + red[2 * TM_SMALL + q] +
                                                  red[3 * TM_SMALL + q];
    }
}

// Inference on blocks of 256 symbols: as k_tmlp_fwd<M, false>, but each wave
// owns two 32-symbol row tiles (symbols 64 w .. 64 w + 63), so every W1
// operand read from LDS feeds two matrix-core products.
constexpr int TI_ROWS = 2 * TM_ROWS;        // symbols per workgroup
constexpr int TI_WIN = TI_ROWS + 4 * TM_R;  // window
constexpr int TI_QW = TI_ROWS + 2 * TM_R;   // pair-table positions

template <int M>
__global__ __launch_bounds__(TM_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_tmlp_infer(TmlpArgs a) {
    __shared__ cx<float> xw[M * TI_WIN];
    __shared__ cx<float> qt2[2 * TI_QW];
    __shared__ __align__(16) unsigned short w1s[TM_H * TM_LDW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, cl = lane & 31, hl = lane >> 5;
    const int row0 = blockIdx.x * TI_ROWS;
    const int seq = row0 / a.S, s0 = row0 % a.S, md = seq % M;
    tm_load_window<M, TI_WIN>(xw, a.x, seq, s0, a.S);
    uint4 pw0, pw1, pw2, pw3;
    auto w1_src = [&](int c, int q) {
        const int idx = threadIdx.x + TM_NT * q;
        return reinterpret_cast<const uint4*>(a.w1b + (size_t)(idx >> 3) * TM_KF + TM_KC * c + 8 * (idx & 7));
    };
    auto w1_dst = [&](int q) {
        const int idx = threadIdx.x + TM_NT * q;
        return reinterpret_cast<uint4*>(w1s + (idx >> 3) * TM_LDW + 8 * (idx & 7));
    };
    auto fetch = [&](int c) {
        pw0 = *w1_src(c, 0);
        pw1 = *w1_src(c, 1);
        pw2 = *w1_src(c, 2);
        pw3 = *w1_src(c, 3);
    };
    auto stage = [&]() {
        *w1_dst(0) = pw0;
        *w1_dst(1) = pw1;
        *w1_dst(2) = pw2;
        *w1_dst(3) = pw3;
    };
    auto chunk_m = [](int c) { return (TM_KC / 2) * c / TM_W - TM_R; };
    fetch(0);
    stage();
    __syncthreads();  // window loaded
    tm_build_pairs2<M, TI_WIN, TI_QW>(qt2, xw, chunk_m(0));
    const cx<float>* xm = xw + md * TI_WIN;
    tm_f16 acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[rt][t] = tm_f16{};
    const int c_sym0 = 2 * TM_R + 64 * w + cl;  // window position of the lane's symbol in row tile 0
    constexpr int NCH = TM_KF / TM_KC;
    __syncthreads();
    for (int c = 0; c < NCH; ++c) {
        if (c + 1 < NCH) fetch(c + 1);
        const int m_lo = chunk_m(c);
        const int t0 = (TM_KC / 2) * c + 4 * hl;
        const int m0 = t0 / TM_W, n0 = t0 - TM_W * m0;
        // row tile rt sits 32 rt window positions further on
        const int qbase = (m0 - TM_R - m_lo) * TI_QW + c_sym0 + n0 - 2 * TM_R;
        const int xbase = c_sym0 + m0 - TM_R;
        cx<float> u_lo[2], u_hi[2];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
            u_lo[rt] = xm[xbase + 32 * rt];
            u_hi[rt] = xm[xbase + 32 * rt + 1];
        }
#pragma unroll
        for (int ks = 0; ks < TM_KC / 16; ++ks) {
            tm_bf8 A[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) {
                uint4 av;
                unsigned* ap = reinterpret_cast<unsigned*>(&av);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // branch-free: the padding slots t >= 1089 read in-bounds LDS and are zeroed
                    const int d = 8 * ks + q;
                    const bool carry = n0 + d >= TM_W;
                    const cx<float> qv = qt2[qbase + 32 * rt + d + (carry ? TI_QW - TM_W : 0)];
                    const cx<float> u = carry ? u_hi[rt] : u_lo[rt];
                    const unsigned v = tm_pack2(u.x * qv.x - u.y * qv.y, u.x * qv.y + u.y * qv.x);
                    ap[q] = t0 + d < TM_T ? v : 0u;
                }
                A[rt] = tm_frag(av);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const tm_bf8 B =
                    tm_frag(*reinterpret_cast<const uint4*>(w1s + (32 * t + cl) * TM_LDW + 16 * ks + 8 * hl));
                acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0], B, acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1], B, acc[1][t], 0, 0, 0);
            }
        }
        if (c + 1 < NCH) {
            __syncthreads();  // every wave is done with chunk c
            stage();
            tm_build_pairs2<M, TI_WIN, TI_QW>(qt2, xw, chunk_m(c + 1));
            __syncthreads();
        }
    }
    float b1v[4], w20[4], w21[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        b1v[t] = a.b1[32 * t + cl];
        w20[t] = a.w2[32 * t + cl];
        w21[t] = a.w2[TM_H + 32 * t + cl];
    }
    const float b20 = a.b2[0], b21 = a.b2[1];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        // lane holds pre-activations of hidden j = 32 t + cl for the symbols
        // r(v) = 8 (v / 4) + 4 hl + v % 4 of row tile rt
        float o0[16], o1[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            float s0v = 0.0f, s1v = 0.0f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float h = fmaxf(acc[rt][t][v] + b1v[t], 0.0f);
                s0v = fmaf(w20[t], h, s0v);
                s1v = fmaf(w21[t], h, s1v);
            }
            o0[v] = s0v;
            o1[v] = s1v;
        }
#pragma unroll
        for (int off = 1; off < 32; off <<= 1)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                o0[v] += __shfl_xor(o0[v], off, 64);
                o1[v] += __shfl_xor(o1[v], off, 64);
            }
        if (cl < 16) {
            const int v = cl, r = 8 * (v >> 2) + 4 * hl + (v & 3);
            float y0 = 0.0f, y1 = 0.0f;
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (u == v) {
                    y0 = o0[u];
                    y1 = o1[u];
                }
            const int pos = 64 * w + 32 * rt + r;  // symbol of the block
            const cx<float> xv = xm[2 * TM_R + pos];
            a.out[(long long)row0 + pos] = mk<float>(xv.x + y0 + b20, xv.y + y1 + b21);
        }
    }
}

// Weight gradient dW1 = dpre f^T on the matrix cores: the reduction runs over
// symbols.  blockIdx.x picks 4 blocks of 32 triplets (one per wave), blockIdx.y
// a contiguous range of 128-symbol blocks.  A operand: dpre rows (hidden
// unit, 8 consecutive symbols per lane); B operand: the lane's triplet over
// the same 8 symbols, real and imaginary parts as two column tiles.
template <int M>
__global__ __launch_bounds__(TM_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_tmlp_dw1(TmlpArgs a,
                                                                                            int splits) {
    __shared__ cx<float> xw[M * TM_WIN];
    __shared__ cx<float> qt[TM_W * TM_QW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, cl = lane & 31, hl = lane >> 5;
    const int tb = 4 * blockIdx.x + w;  // triplet block of this wave
    const int t = 32 * tb + cl;
    const bool live = t < TM_T;
    const int tm = t / TM_W - TM_R, tn = t % TM_W - TM_R;
    const int tmc = live ? tm : 0, tnc = live ? tn : 0;
    const int split = blockIdx.y;
    const int b_lo = (int)((long long)a.blocks * split / splits), b_hi = (int)((long long)a.blocks * (split + 1) / splits);
    tm_f16 accr[4], acci[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) accr[q] = acci[q] = tm_f16{};
    for (int blk = b_lo; blk < b_hi; ++blk) {
        const int row0 = blk * TM_ROWS;
        const int seq = row0 / a.S, s0 = row0 % a.S, md = seq % M;
        __syncthreads();  // the previous block's window and table are no longer read
        tm_load_window<M>(xw, a.x, seq, s0, a.S);
        __syncthreads();
        tm_build_pairs<M>(qt, xw);
        __syncthreads();
        const cx<float>* xm = xw + md * TM_WIN;
        if (tb < TM_TP / 32) {
#pragma unroll 1
            for (int ks = 0; ks < TM_ROWS / 16; ++ks) {
                const int r = 16 * ks + 8 * hl;  // first of the lane's 8 symbols
                // the A operands first: their loads fly while the triplets form
                uint4 av[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    av[q] = *reinterpret_cast<const uint4*>(a.dpre + (size_t)(32 * q + cl) * a.rows + row0 + r);
                uint4 br, bi;
                unsigned* pr = reinterpret_cast<unsigned*>(&br);
                unsigned* pi = reinterpret_cast<unsigned*>(&bi);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    // branch-free: a padding lane (t >= 1089) reads the table
                    // at m = n = 0 and is zeroed
                    const cx<float> T0 = tm_triplet(qt, xm, 2 * TM_R + r + 2 * u, tmc, tnc);
                    const cx<float> T1 = tm_triplet(qt, xm, 2 * TM_R + r + 2 * u + 1, tmc, tnc);
                    pr[u] = live ? tm_pack2(T0.x, T1.x) : 0u;
                    pi[u] = live ? tm_pack2(T0.y, T1.y) : 0u;
                }
                const tm_bf8 Br = tm_frag(br), Bi = tm_frag(bi);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const tm_bf8 A = tm_frag(av[q]);
                    accr[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Br, accr[q], 0, 0, 0);
                    acci[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bi, acci[q], 0, 0, 0);
                }
            }
        }
    }
    if (tb >= TM_TP / 32) return;
    float* P = a.part + (size_t)split * TM_H * TM_KF;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int j = 32 * q + 8 * (v >> 2) + 4 * hl + (v & 3);
            *reinterpret_cast<float2*>(P + (size_t)j * TM_KF + 2 * t) = make_float2(accr[q][v], acci[q][v]);
        }
}

// grad[i] += sum over the splits of the weight-gradient partials (W1 part)
__global__ void k_tmlp_reduce_w1(const float* __restrict__ part, int splits, double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= TM_H * TM_KF) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += (double)part[(size_t)k * TM_H * TM_KF + i];
    grad[i] += s;
}

// one workgroup per small quantity q: sum over the forward blocks; the last
// quantity is the loss
__global__ void k_tmlp_reduce_small(const float* __restrict__ small, int blocks, double* __restrict__ grad,
                                    double* __restrict__ loss) {
    __shared__ double red[TM_NT / 64];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < blocks; b += TM_NT) s += (double)small[(size_t)q * blocks + b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int k = 0; k < TM_NT / 64; ++k) tot += red[k];
        if (q < TM_SMALL - 1)
            grad[TM_H * TM_KF + q] += tot;
        else if (loss)
            *loss += tot;
    }
}

// (1/count) sum |out - target|^2 into *loss
__global__ void k_tmlp_loss(const cx<float>* __restrict__ out, const cx<float>* __restrict__ target, long long n,
                            double inv_count, double* __restrict__ loss) {
    __shared__ double red[TM_NT / 64];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)TM_NT + threadIdx.x; i < n; i += (long long)gridDim.x * TM_NT) {
        const cx<float> d = out[i] - target[i];
        s += (double)d.x * d.x + (double)d.y * d.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int k = 0; k < TM_NT / 64; ++k) tot += red[k];
        atomicAdd(loss, tot * inv_count);
    }
}

}  // namespace focss_hip
