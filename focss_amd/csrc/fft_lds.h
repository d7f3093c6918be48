// Stockham auto-sort FFT held in LDS (replaces the FFTW plans of
// src/focss/field.cpp:530-541).  A tile holds S independent sequences of
// length L interleaved as element (i, s) at lds[pad(i * S + s)], so
// consecutive lanes touch consecutive LDS words in every stage; pad() inserts
// one element every 8 so that strided transposing loads spread over banks.
// Each stage is the radix-R step of the Stockham formulation in Govindaraju
// et al., "High performance discrete Fourier transforms on graphics
// processors" (SC'08): read R values strided by L/R, twiddle by
// W_{Ns R}^{r k}, radix-R DFT, write to expand(j) + r Ns.  Radix-8 stages
// first, then one radix-4 or radix-2 stage for the remainder.
#pragma once
#include <type_traits>
#include "cx.h"

namespace focss_hip {

// Twiddle banks in global memory: tw[off(p) + m] = exp(-2 pi i m / 2^p),
// m < 2^p, p = 1..TW_MAX_LOG.  Rounded once from long double on the host.
constexpr int TW_MAX_LOG = 14;
__host__ __device__ constexpr int tw_offset(int log_len) { return (1 << log_len) - 2; }
constexpr int TW_BANK_SIZE = (1 << (TW_MAX_LOG + 1)) - 2;

template <int L>
struct ilog2 {
    static constexpr int value = 1 + ilog2<L / 2>::value;
};
template <>
struct ilog2<1> {
    static constexpr int value = 0;
};

__host__ __device__ constexpr int pad(int p) { return p + (p >> 3); }
__host__ __device__ constexpr int padded(int n) { return n + (n >> 3) + 1; }

// W_L^m from a full length-L table in LDS
template <typename R>
struct TwFull {
    static constexpr bool kPowers = false;
    const cx<R>* t;
    __device__ __forceinline__ cx<R> get(int m) const { return t[m]; }
};

// W_L^m = hi[m >> LB] * lo[m & (2^LB - 1)] from two short LDS tables
template <typename R, int LB>
struct TwSplit {
    static constexpr bool kPowers = false;
    const cx<R>* lo;
    const cx<R>* hi;
    __device__ __forceinline__ cx<R> get(int m) const { return hi[m >> LB] * lo[m & ((1 << LB) - 1)]; }
};

// The same two tables for the split-stage DFTs, whose stages look up one
// twiddle W per butterfly and form W^r, r < RADIX, in registers
template <typename R, int LB>
struct TwSplitPow : TwSplit<R, LB> {
    static constexpr bool kPowers = true;
};

// Fill the LDS twiddle tables for length L from the global bank.
template <typename R, int L, int NT>
__device__ __forceinline__ void load_tw_full(cx<R>* t, const cx<R>* __restrict__ bank) {
    const cx<R>* src = bank + tw_offset(ilog2<L>::value);
    for (int i = threadIdx.x; i < L; i += NT) t[i] = src[i];
}
template <typename R, int L, int LB, int NT>
__device__ __forceinline__ void load_tw_split(cx<R>* lo, cx<R>* hi, const cx<R>* __restrict__ bank) {
    const cx<R>* src = bank + tw_offset(ilog2<L>::value);
    for (int i = threadIdx.x; i < (1 << LB); i += NT) lo[i] = src[i];
    for (int i = threadIdx.x; i < L / (1 << LB); i += NT) hi[i] = src[i << LB];
}

// radix-2/4/8/16 DFT in registers, exponent sign SIGN (-1 forward, +1 inverse)
template <typename R, int RADIX, int SIGN>
struct butterfly;

template <typename R, int SIGN>
struct butterfly<R, 2, SIGN> {
    static __device__ __forceinline__ void run(cx<R>* v) {
        cx<R> a = v[0] + v[1];
        cx<R> b = v[0] - v[1];
        v[0] = a;
        v[1] = b;
    }
};

template <typename R, int SIGN>
struct butterfly<R, 4, SIGN> {
    static __device__ __forceinline__ void run(cx<R>* v) {
        cx<R> a0 = v[0] + v[2];
        cx<R> a1 = v[0] - v[2];
        cx<R> b0 = v[1] + v[3];
        cx<R> b1 = mul_si<R, SIGN>(v[1] - v[3]);
        v[0] = a0 + b0;
        v[2] = a0 - b0;
        v[1] = a1 + b1;
        v[3] = a1 - b1;
    }
};

template <typename R, int SIGN>
struct butterfly<R, 8, SIGN> {
    static __device__ __forceinline__ void run(cx<R>* v) {
        const R h = (R)0.70710678118654752440084436210484904;
        cx<R> a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a[k] = v[k] + v[k + 4];
            b[k] = v[k] - v[k + 4];
        }
        // b_k *= W8^k, W8 = exp(SIGN 2 pi i / 8)
        b[1] = mk<R>((b[1].x - SIGN * b[1].y) * h, (b[1].y + SIGN * b[1].x) * h);
        b[2] = mul_si<R, SIGN>(b[2]);
        b[3] = mk<R>((-b[3].x - SIGN * b[3].y) * h, (-b[3].y + SIGN * b[3].x) * h);
        butterfly<R, 4, SIGN>::run(a);
        butterfly<R, 4, SIGN>::run(b);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[2 * k] = a[k];
            v[2 * k + 1] = b[k];
        }
    }
};

template <typename R, int SIGN>
struct butterfly<R, 16, SIGN> {
    // 16 = 4 x 4: length-4 DFTs over n1 (stride 4), twiddles W16^{n2 k1},
    // length-4 DFTs over n2; output k1 + 4 k2
    static __device__ __forceinline__ void run(cx<R>* v) {
        const R h = (R)0.70710678118654752440084436210484904;
        const R c1 = (R)0.92387953251128675612818318939678829;  // cos(pi/8)
        const R s1 = (R)0.38268343236508977172845998403039887;  // sin(pi/8)
        cx<R> a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[n2 + 4 * n1];
            butterfly<R, 4, SIGN>::run(a[n2]);
        }
        // a[n2][k1] *= W16^{n2 k1}, W16 = exp(SIGN 2 pi i / 16)
        const cx<R> w1 = mk<R>(c1, SIGN * s1), w2 = mk<R>(h, SIGN * h), w3 = mk<R>(s1, SIGN * c1);
        a[1][1] = a[1][1] * w1;
        a[1][2] = a[1][2] * w2;
        a[1][3] = a[1][3] * w3;
        a[2][1] = a[2][1] * w2;
        a[2][2] = mul_si<R, SIGN>(a[2][2]);
        a[2][3] = mul_si<R, SIGN>(a[2][3] * w2);
        a[3][1] = a[3][1] * w3;
        a[3][2] = mul_si<R, SIGN>(a[3][2] * w2);
        a[3][3] = (a[3][3] * w1) * (R)-1;  // W16^9 = -W16
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cx<R> b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            butterfly<R, 4, SIGN>::run(b);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
        }
    }
};

// One Stockham stage over S interleaved sequences of length L with NT
// threads, then the remaining stages.
template <typename R, int L, int S, int NT, int SIGN, int NS, typename TW>
__device__ __forceinline__ void fft_stages(cx<R>* __restrict__ buf, const TW& tw) {
    constexpr int RADIX = (L / NS >= 8) ? 8 : (L / NS);
    constexpr int NB = L / RADIX;
    constexpr int TOTAL = NB * S;
    constexpr int PER = (TOTAL + NT - 1) / NT;
    constexpr int TWS = L / (NS * RADIX);
    const int tid = threadIdx.x;
    cx<R> v[PER][RADIX];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = tid + b * NT;
        if (TOTAL % NT == 0 || id < TOTAL) {
            const int s = id % S;
            const int j = id / S;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[b][r] = buf[pad((j + r * NB) * S + s)];
            if (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < RADIX; ++r) {
                    cx<R> w = tw.get(r * k * TWS);
                    v[b][r] = v[b][r] * (SIGN < 0 ? w : conj(w));
                }
            }
            butterfly<R, RADIX, SIGN>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = tid + b * NT;
        if (TOTAL % NT == 0 || id < TOTAL) {
            const int s = id % S;
            const int j = id / S;
            const int k = j % NS;
            const int base = (j / NS) * NS * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) buf[pad((base + r * NS) * S + s)] = v[b][r];
        }
    }
    __syncthreads();
    if constexpr (NS * RADIX < L) fft_stages<R, L, S, NT, SIGN, NS * RADIX, TW>(buf, tw);
}

// Unscaled DFT of S interleaved length-L sequences in LDS (natural order in
// and out).  The caller's writes must be visible (barrier) before the call;
// the call ends with a barrier.
template <typename R, int L, int S, int NT, int SIGN, typename TW>
__device__ __forceinline__ void fft_lds(cx<R>* __restrict__ buf, const TW& tw) {
    if constexpr (L > 1) fft_stages<R, L, S, NT, SIGN, 1, TW>(buf, tw);
}

// Twiddle tables of the split-stage DFTs of length L: FP64 from the two
// short tables with powers formed in registers (little LDS, two lookups
// per butterfly); FP32 from the full table (powers of a rounded complex64
// W drift: 1.1e-3 against 1.7e-4 field error over 1500 split steps)
template <typename R, int L>
struct PassTw {
    static constexpr bool kSplit = sizeof(R) == 8;
    static constexpr int LB = (ilog2<L>::value + 1) / 2;
    static constexpr int A = kSplit ? (1 << LB) : L;  // sizes of the two LDS arrays
    static constexpr int B = kSplit ? (L >> LB) : 1;
    using type = std::conditional_t<kSplit, TwSplitPow<R, LB>, TwFull<R>>;
    template <int NT>
    static __device__ __forceinline__ type load(cx<R>* a, cx<R>* b, const cx<R>* __restrict__ bank) {
        if constexpr (kSplit) {
            load_tw_split<R, L, LB, NT>(a, b, bank);
            return type{{a, b}};
        } else {
            load_tw_full<R, L, NT>(a, bank);
            return type{a};
        }
    }
    // The same in two halves: fetch() issues the global loads into
    // registers, store() writes them to LDS.  A kernel issues its field
    // loads between the two, so that the table's latency and the field's
    // overlap instead of adding up.
    template <int NT>
    struct Staged {
        static constexpr int PA = (A + NT - 1) / NT, PB = (B + NT - 1) / NT;
        cx<R> a[PA], b[PB];
    };
    template <int NT>
    static __device__ __forceinline__ Staged<NT> fetch(const cx<R>* __restrict__ bank) {
        const cx<R>* src = bank + tw_offset(ilog2<L>::value);
        Staged<NT> s;
#pragma unroll
        for (int k = 0; k < Staged<NT>::PA; ++k) {
            const int i = threadIdx.x + k * NT;
            if (A % NT == 0 || i < A) s.a[k] = src[i];
        }
        if constexpr (kSplit) {
#pragma unroll
            for (int k = 0; k < Staged<NT>::PB; ++k) {
                const int i = threadIdx.x + k * NT;
                if (B % NT == 0 || i < B) s.b[k] = src[i << LB];
            }
        }
        return s;
    }
    template <int NT>
    static __device__ __forceinline__ type store(cx<R>* a, cx<R>* b, const Staged<NT>& s) {
#pragma unroll
        for (int k = 0; k < Staged<NT>::PA; ++k) {
            const int i = threadIdx.x + k * NT;
            if (A % NT == 0 || i < A) a[i] = s.a[k];
        }
        if constexpr (kSplit) {
#pragma unroll
            for (int k = 0; k < Staged<NT>::PB; ++k) {
                const int i = threadIdx.x + k * NT;
                if (B % NT == 0 || i < B) b[i] = s.b[k];
            }
            return type{{a, b}};
        } else {
            return type{a};
        }
    }
};

// ------------------------------------------------------------------
// Split-stage DFTs for the two-pass kernels.  A length-L DFT (L >= 64) runs
// the mixed-radix schedule P(L) = (8, 2^(log2 L mod 3), 8, ..., 8), or its
// reverse; both begin and end with a radix-8 stage.  The first stage reads
// x[j + r L/8] and the last one writes y[j + r L/8] (r < 8, j < L/8), so a
// thread that owns item j holds exactly the elements the neighbouring
// operation needs: the first stage is fed from HBM, the last one feeds
// either HBM or the pointwise operator and the next DFT's first stage
// (which must use the reversed schedule), with no LDS round trip.
// Stage twiddles W_{NS R}^{r k} = W_L^{r k L/(NS R)}: one lookup of
// W_L^{k L/(NS R)} per butterfly (TwSplitPow: two short LDS tables), its
// powers by products of depth <= 3 (radix 8) or 4 (radix 16).
// ------------------------------------------------------------------
template <int L, bool REV>
struct Sched {
    static constexpr int P = ilog2<L>::value;
    static constexpr int REM = P % 3;
    // log2 L = 3k + 1 with k >= 3: the remainder joins a radix-8 stage as one
    // radix-16 middle stage, (8, 16, 8, ...), one LDS round trip fewer than
    // (8, 2, 8, 8, ...)
    static constexpr bool R16 = REM == 1 && P >= 10;
    static constexpr int N = R16 ? (P - 1) / 3 : P / 3 + (REM ? 1 : 0);
    static constexpr int fwd_radix(int i) { return i == 1 ? (R16 ? 16 : (REM ? (1 << REM) : 8)) : 8; }
    static constexpr int radix(int i) { return REV ? fwd_radix(N - 1 - i) : fwd_radix(i); }
    static constexpr int ns(int i) {
        int p = 1;
        for (int k = 0; k < i; ++k) p *= radix(k);
        return p;
    }
};

// w[r] = W_L^{r k STRIDE} (SIGN -1) or its conjugate (SIGN +1), 1 <= r < RADIX,
// from the table tw: every power looked up, or (kPowers) one lookup and
// products
template <typename R, int RADIX, int SIGN, int STRIDE, typename TW>
__device__ __forceinline__ void stage_twiddles(const TW& tw, int k, cx<R>* w) {
    if constexpr (TW::kPowers) {
        cx<R> w1 = tw.get(k * STRIDE);
        if constexpr (SIGN > 0) w1 = conj(w1);
        w[1] = w1;
        if constexpr (RADIX >= 4) {
            w[2] = w1 * w1;
            w[3] = w[2] * w1;
        }
        if constexpr (RADIX >= 8) {
            w[4] = w[2] * w[2];
            w[5] = w[4] * w1;
            w[6] = w[4] * w[2];
            w[7] = w[4] * w[3];
        }
        if constexpr (RADIX >= 16) {
            w[8] = w[4] * w[4];
#pragma unroll
            for (int r = 9; r < 16; ++r) w[r] = w[8] * w[r - 8];
        }
    } else {
#pragma unroll
        for (int r = 1; r < RADIX; ++r) {
            w[r] = tw.get(r * k * STRIDE);
            if constexpr (SIGN > 0) w[r] = conj(w[r]);
        }
    }
}

// One Stockham stage (radix RADIX, NS = product of earlier radices) over S
// interleaved sequences in LDS, NT threads; barriers before the writes and
// after them.  tw gives W_L^m.
template <typename R, int L, int S, int NT, int SIGN, int RADIX, int NS, typename TW>
__device__ __forceinline__ void stockham_stage(cx<R>* __restrict__ buf, const TW& tw) {
    constexpr int NB = L / RADIX;
    constexpr int TOTAL = NB * S;
    constexpr int PER = (TOTAL + NT - 1) / NT;
    const int tid = threadIdx.x;
    cx<R> v[PER][RADIX];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = tid + b * NT;
        if (TOTAL % NT == 0 || id < TOTAL) {
            const int s = id % S;
            const int j = id / S;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[b][r] = buf[pad((j + r * NB) * S + s)];
            if constexpr (NS > 1) {
                cx<R> w[RADIX];
                stage_twiddles<R, RADIX, SIGN, L / (NS * RADIX)>(tw, j % NS, w);
#pragma unroll
                for (int r = 1; r < RADIX; ++r) v[b][r] = v[b][r] * w[r];
            }
            butterfly<R, RADIX, SIGN>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int id = tid + b * NT;
        if (TOTAL % NT == 0 || id < TOTAL) {
            const int s = id % S;
            const int j = id / S;
            const int k = j % NS;
            const int base = (j / NS) * NS * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) buf[pad((base + r * NS) * S + s)] = v[b][r];
        }
    }
    __syncthreads();
}

// stages 1 .. N-2 of the schedule (the first and last run in registers)
template <typename R, int L, int S, int NT, int SIGN, bool REV, int I = 1, typename TW>
__device__ __forceinline__ void middle_stages(cx<R>* __restrict__ buf, const TW& tw) {
    using SC = Sched<L, REV>;
    if constexpr (I < SC::N - 1) {
        stockham_stage<R, L, S, NT, SIGN, SC::radix(I), SC::ns(I)>(buf, tw);
        middle_stages<R, L, S, NT, SIGN, REV, I + 1>(buf, tw);
    }
}

// first stage: radix-8 DFT of x[j + r L/8] held in v, written to the LDS
// positions 8 j + r of sequence s (no barrier)
template <typename R, int L, int S, int SIGN>
__device__ __forceinline__ void first_stage_store(cx<R>* __restrict__ buf, cx<R>* v, int s, int j) {
    butterfly<R, 8, SIGN>::run(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[pad((j * 8 + r) * S + s)] = v[r];
}

// last stage: reads x[j + r L/8] of sequence s, twiddles by W_L^{r j},
// radix-8 DFT; v[r] is then output element j + r L/8
template <typename R, int L, int S, int SIGN, typename TW>
__device__ __forceinline__ void last_stage_load(const cx<R>* __restrict__ buf, const TW& tw, cx<R>* v, int s,
                                                int j) {
    constexpr int Q = L / 8;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = buf[pad((j + r * Q) * S + s)];
    cx<R> w[8];
    stage_twiddles<R, 8, SIGN, 1>(tw, j, w);
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = v[r] * w[r];
    butterfly<R, 8, SIGN>::run(v);
}

}  // namespace focss_hip
