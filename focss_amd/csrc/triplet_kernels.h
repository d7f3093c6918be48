// Perturbation-triplet equalizer kernels (src/focss/equalizer/sino.cpp,
// src/focss/equalizer/wdm_sino.cpp).
//
//  k_sino_cubic  : cubic(k) = sum_{|m|+|n|<=R} W(m,n) A_mode(k+m)
//                             sum_p A_p(k+n) conj(A_p(k+m+n))      (sino.cpp:203-220)
//                  one output sample per thread, the field window
//                  [k0-R, k0+K+R) of every mode staged in LDS; weights are
//                  read in diamond order with wave-uniform indices (scalar
//                  loads).
//  k_sino_features: the folded symmetric feature rows of SINO::train
//                  (sino.cpp:67-103), column-major rows x n_features.
//  k_gram_mfma   : G = F^H F over a column-major complex matrix on the FP64
//                  matrix cores, upper 32x32 tiles split over row slices.
//  k_col_sums    : sum_r F(r, j), sum_r conj(F(r, j)) y(r), sum_r y(r).
//  k_wdm_ppe     : WdmPPE::spm + WdmPPE::xpm over 6 envelopes (wdm_sino.cpp:256-295).
//  k_wdmppe_features: the feature rows of WdmPPE::train (wdm_sino.cpp:29-81).
#pragma once
#include "cx.h"

namespace focss_hip {

constexpr int TRIP_NT = 256;

// diamond index of (m, n), |m| + |n| <= R, m outer then n inner
__host__ __device__ inline int diamond_offset(int m, int R) {
    // sum over m' in [-R, m) of (2 (R - |m'|) + 1)
    int off = 0;
    for (int mm = -R; mm < m; ++mm) off += 2 * (R - (mm < 0 ? -mm : mm)) + 1;
    return off;
}

template <typename R>
__global__ __launch_bounds__(TRIP_NT) void k_sino_cubic(const cx<R>* __restrict__ in, cx<R>* __restrict__ out,
                                                        const cx<R>* __restrict__ wdiam, int radius, int modes,
                                                        int samples) {
    extern __shared__ unsigned char smem_raw[];
    cx<R>* win = reinterpret_cast<cx<R>*>(smem_raw);
    const int W = TRIP_NT + 2 * radius;
    const int k0 = blockIdx.x * TRIP_NT;
    for (int i = threadIdx.x; i < modes * W; i += TRIP_NT) {
        int p = i / W, j = i % W;
        int k = k0 - radius + j;
        win[i] = (k >= 0 && k < samples) ? in[(long long)p * samples + k] : mk<R>(0, 0);
    }
    __syncthreads();
    const int k = k0 + threadIdx.x;
    if (k >= samples) return;
    const bool interior = k >= radius && k < samples - radius;
    const int j = threadIdx.x + radius;  // window index of sample k
    for (int mode = 0; mode < modes; ++mode) {
        cx<R> acc = mk<R>(0, 0);
        if (interior) {
            int idx = 0;
            for (int m = -radius; m <= radius; ++m) {
                const int b = radius - (m < 0 ? -m : m);
                cx<R> inner = mk<R>(0, 0);
                for (int n = -b; n <= b; ++n, ++idx) {
                    cx<R> t = mk<R>(0, 0);
                    for (int p = 0; p < modes; ++p)
                        t = t + win[p * W + j + n] * conj(win[p * W + j + m + n]);
                    inner = inner + wdiam[idx] * t;
                }
                acc = acc + inner * win[mode * W + j + m];
            }
        }
        out[(long long)mode * samples + k] = acc;
    }
}

// SINO::term (sino.cpp:161-171) on a mode-major field in global memory
__device__ __forceinline__ cx<double> sino_term(const cx<double>* f, int modes, int samples, int mode, int k, int m,
                                                int n) {
    cx<double> t = mk<double>(0, 0);
    for (int p = 0; p < modes; ++p)
        t = t + f[(long long)p * samples + k + n] * conj(f[(long long)p * samples + k + m + n]);
    return f[(long long)mode * samples + k + m] * t;
}

// one thread per row; features column-major F[row + rows * j]
__global__ void k_sino_features(const cx<double>* __restrict__ tx, const cx<double>* __restrict__ rx,
                                cx<double>* __restrict__ F, cx<double>* __restrict__ y, int radius, int type,
                                cx<double> lin_w, int modes, int samples) {
    const int per_mode = samples - 2 * radius;
    const long long rows = (long long)per_mode * modes;
    for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < rows;
         row += (long long)gridDim.x * blockDim.x) {
        const int mode = (int)(row / per_mode);
        const int s = (int)(row % per_mode);
        const int k = s + radius;
        cx<double> r = rx[(long long)mode * samples + k];
        if (type == 1) r = r - tx[(long long)mode * samples + k] * lin_w;
        y[row] = r;
        int index = 0;
        const int mb = radius / 2;
        for (int m = -mb; m <= mb; ++m) {
            const int am = m < 0 ? -m : m;
            for (int n = am; n <= radius - am; ++n) {
                cx<double> v;
                if (m == 0 && n == 0) {
                    v = sino_term(tx, modes, samples, mode, k, 0, 0);
                } else if (m == n) {
                    v = sino_term(tx, modes, samples, mode, k, m, n) + sino_term(tx, modes, samples, mode, k, -m, -n);
                } else if (m == -n) {
                    v = sino_term(tx, modes, samples, mode, k, m, n) + sino_term(tx, modes, samples, mode, k, n, m);
                } else {
                    v = sino_term(tx, modes, samples, mode, k, m, n) + sino_term(tx, modes, samples, mode, k, n, m) +
                        sino_term(tx, modes, samples, mode, k, -m, -n) + sino_term(tx, modes, samples, mode, k, -n, -m);
                }
                F[row + rows * index++] = v;
            }
        }
        if (type == 0) F[row + rows * index] = tx[(long long)mode * samples + k];
    }
}

// G += F^H F on the FP64 matrix cores (v_mfma_f64_16x16x4f64), upper tiles
// only (G is Hermitian; the caller mirrors the lower triangle).  With F = A +
// i B: Re G = A^T A + B^T B, Im G = A^T B - B^T A.  A workgroup owns one
// 32 x 32 complex tile (ti <= tj, blockIdx.x over the upper pairs) and the
// rows of one K slice (blockIdx.y of gridDim.y); wave w computes the real
// 32 x 32 block P^T Q of (P, Q) = (A, A), (B, B), (A, B), (B, A) as 2 x 2
// MFMA tiles from row chunks staged in LDS, and adds it with its sign into
// G (FP64 atomics, G zeroed by the caller).
// MFMA lane layout (FP64 16x16x4): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator v holds D[4 v + l / 16][l % 16] (one row per
// register, unlike the 4-row groups of the FP16 16x16 instructions).
constexpr int GM_T = 32;   // complex columns per tile side
constexpr int GM_KC = 16;  // rows per LDS chunk
typedef double gm_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gram_mfma(const cx<double>* __restrict__ F, cx<double>* __restrict__ G,
                                                   long long rows, int nf, int tiles_side) {
    __shared__ cx<double> fi[GM_KC][GM_T + 1];
    __shared__ cx<double> fj[GM_KC][GM_T + 1];
    // upper-triangular pair (ti, tj), ti <= tj, from the linear index
    int p = blockIdx.x, ti = 0;
    while (p >= tiles_side - ti) {
        p -= tiles_side - ti;
        ++ti;
    }
    const int tj = ti + p;
    const int i0 = ti * GM_T, j0 = tj * GM_T;
    const long long per = (rows + gridDim.y - 1) / gridDim.y;
    const long long r_begin = blockIdx.y * per;
    const long long r_end = r_begin + per < rows ? r_begin + per : rows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool p_im = wave == 1 || wave == 3;  // P = B for waves 1 (B,B) and 3 (B,A)
    const bool q_im = wave == 1 || wave == 2;  // Q = B for waves 1 (B,B) and 2 (A,B)
    gm_d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = gm_d4{0, 0, 0, 0};
    for (long long r0 = r_begin; r0 < r_end; r0 += GM_KC) {
        // 2 x 512 complex values: consecutive threads take consecutive rows of one column
        for (int e = threadIdx.x; e < 2 * GM_KC * GM_T; e += 256) {
            const int which = e / (GM_KC * GM_T), f = e % (GM_KC * GM_T);
            const int c = f / GM_KC, rr = f % GM_KC;
            const long long r = r0 + rr;
            const int col = (which ? j0 : i0) + c;
            const cx<double> v = (r < r_end && col < nf) ? F[r + rows * col] : mk<double>(0, 0);
            if (which)
                fj[rr][c] = v;
            else
                fi[rr][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GM_KC / 4; ++ks) {
            const int k = ks * 4 + (lane >> 4);
            double av[2], bv[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const cx<double> x = fi[k][t * 16 + (lane & 15)];
                const cx<double> y = fj[k][t * 16 + (lane & 15)];
                av[t] = p_im ? x.y : x.x;
                bv[t] = q_im ? y.y : y.x;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // waves 0, 1 add to Re G; wave 2 adds, wave 3 subtracts, Im G
    const double sign = wave == 3 ? -1.0 : 1.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + a * 16 + 4 * v + (lane >> 4);
                const int j = j0 + b * 16 + (lane & 15);
                if (i < nf && j < nf) {
                    double* g = reinterpret_cast<double*>(G + (long long)i * nf + j);
                    atomicAdd(g + (wave >= 2 ? 1 : 0), sign * acc[a][b][v]);
                }
            }
}

// per column j: sum_r F(r, j) and sum_r conj(F(r, j)) y(r); column nf holds sum_r y(r)
__global__ __launch_bounds__(256) void k_col_sums(const cx<double>* __restrict__ F, const cx<double>* __restrict__ y,
                                                  long long rows, int nf, cx<double>* __restrict__ sums,
                                                  cx<double>* __restrict__ fy) {
    __shared__ cx<double> red_s[256], red_f[256];
    const int j = blockIdx.x;
    cx<double> s = mk<double>(0, 0), f = mk<double>(0, 0);
    for (long long r = threadIdx.x; r < rows; r += 256) {
        cx<double> v = j < nf ? F[r + rows * j] : mk<double>(1, 0);
        s = s + (j < nf ? v : y[r]);
        if (j < nf) f = f + conj(v) * y[r];
    }
    red_s[threadIdx.x] = s;
    red_f[threadIdx.x] = f;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red_s[threadIdx.x] = red_s[threadIdx.x] + red_s[threadIdx.x + o];
            red_f[threadIdx.x] = red_f[threadIdx.x] + red_f[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[j] = red_s[0];
        if (j < nf) fy[j] = red_f[0];
    }
}

// WdmPPE: A6 row-major [6][samples], rows = ch * 2 + pol, COI rows 2..3.
// spm: n outer, m inner (wdm_sino.cpp:261-272); xpm: m outer, n inner with
// T1(m, n) + T3(m, -n) (wdm_sino.cpp:281-293).  Diamond-packed C (n outer) and
// D (m outer).
template <typename R>
__device__ __forceinline__ cx<R> at6(const cx<R>* A6, int samples, int r, int k) {
    return A6[(long long)r * samples + k];
}

// WdmPPE::T (wdm_sino.cpp:155-164): A_{2+pol}(k+m) sum_{p=2,3} A_p(k+n) conj(A_p(k+m+n))
template <typename R>
__device__ __forceinline__ cx<R> ppe_spm_term(const cx<R>* A6, int samples, int pol, int k, int m, int n) {
    cx<R> t = at6(A6, samples, 2, k + n) * conj(at6(A6, samples, 2, k + m + n)) +
              at6(A6, samples, 3, k + n) * conj(at6(A6, samples, 3, k + m + n));
    return at6(A6, samples, 2 + pol, k + m) * t;
}

// WdmPPE::T1(m, n) + WdmPPE::T3(m, -n) (wdm_sino.cpp:166-200): the XPM terms
// of the channels below (rows 0..1) and above (rows 4..5) the COI
template <typename R>
__device__ __forceinline__ cx<R> ppe_xpm_term(const cx<R>* A6, int samples, int pol, int k, int m, int n) {
    const int a = 2 + pol, o = 3 - pol;      // this / other COI polarisation
    const int c0 = pol, c1 = 1 - pol;        // channel 0 rows
    const int e0 = 4 + pol, e1 = 5 - pol;    // channel 2 rows
    cx<R> t1 = (at6(A6, samples, a, k + m) * at6(A6, samples, c0, k + n) * conj(at6(A6, samples, c0, k + m + n))) *
                   (R)2 +
               at6(A6, samples, a, k + m) * at6(A6, samples, c1, k + n) * conj(at6(A6, samples, c1, k + m + n)) +
               at6(A6, samples, o, k + m) * at6(A6, samples, c0, k + n) * conj(at6(A6, samples, c1, k + m + n));
    const int nn = -n;
    cx<R> t3 = (at6(A6, samples, a, k + m) * at6(A6, samples, e0, k + nn) *
                conj(at6(A6, samples, e0, k + m + nn))) * (R)2 +
               at6(A6, samples, a, k + m) * at6(A6, samples, e1, k + nn) * conj(at6(A6, samples, e1, k + m + nn)) +
               at6(A6, samples, o, k + m) * at6(A6, samples, e0, k + nn) * conj(at6(A6, samples, e1, k + m + nn));
    return t1 + t3;
}

template <typename R>
__global__ __launch_bounds__(TRIP_NT) void k_wdm_ppe(const cx<R>* __restrict__ A6, const cx<R>* __restrict__ cdiam,
                                                     int c_radius, const cx<R>* __restrict__ ddiam, int d_radius,
                                                     int samples, cx<R>* __restrict__ spm, cx<R>* __restrict__ xpm) {
    const int k = blockIdx.x * TRIP_NT + threadIdx.x;
    if (k >= samples) return;
    for (int pol = 0; pol < 2; ++pol) {
        cx<R> acc = mk<R>(0, 0);
        if (c_radius >= 0 && k >= c_radius && k < samples - c_radius) {
            int idx = 0;
            for (int n = -c_radius; n <= c_radius; ++n) {
                const int b = c_radius - (n < 0 ? -n : n);
                for (int m = -b; m <= b; ++m, ++idx)
                    acc = acc + cdiam[idx] * ppe_spm_term(A6, samples, pol, k, m, n);
            }
        }
        spm[(long long)pol * samples + k] = acc;
        cx<R> acx = mk<R>(0, 0);
        if (d_radius >= 0 && k >= d_radius && k < samples - d_radius) {
            int idx = 0;
            for (int m = -d_radius; m <= d_radius; ++m) {
                const int b = d_radius - (m < 0 ? -m : m);
                for (int n = -b; n <= b; ++n, ++idx) acx = acx + ddiam[idx] * ppe_xpm_term(A6, samples, pol, k, m, n);
            }
        }
        xpm[(long long)pol * samples + k] = acx;
    }
}

// WdmPPE::train features (wdm_sino.cpp:39-81), one thread per row; rows
// pol-major (row = pol * (samples - 2r) + k - r, r = max(c, d)), F
// column-major [row + rows * j]: the folded SPM set (the fold of SINO::train),
// the XPM diamond (m outer), the COI sample; y = rx rows 2..3.
__global__ void k_wdmppe_features(const cx<double>* __restrict__ tx6, const cx<double>* __restrict__ rx6,
                                  cx<double>* __restrict__ F, cx<double>* __restrict__ y, int c_radius,
                                  int d_radius, int samples) {
    const int r = c_radius > d_radius ? c_radius : d_radius;
    const int per_pol = samples - 2 * r;
    const long long rows = 2LL * per_pol;
    for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < rows;
         row += (long long)gridDim.x * blockDim.x) {
        const int pol = (int)(row / per_pol);
        const int k = (int)(row % per_pol) + r;
        y[row] = rx6[(long long)(2 + pol) * samples + k];
        int index = 0;
        const int mb = c_radius / 2;
        for (int m = -mb; m <= mb; ++m) {
            const int am = m < 0 ? -m : m;
            for (int n = am; n <= c_radius - am; ++n) {
                cx<double> v;
                if (m == 0 && n == 0) {
                    v = ppe_spm_term(tx6, samples, pol, k, 0, 0);
                } else if (m == n) {
                    v = ppe_spm_term(tx6, samples, pol, k, m, n) + ppe_spm_term(tx6, samples, pol, k, -m, -n);
                } else if (m == -n) {
                    v = ppe_spm_term(tx6, samples, pol, k, m, n) + ppe_spm_term(tx6, samples, pol, k, n, m);
                } else {
                    v = ppe_spm_term(tx6, samples, pol, k, m, n) + ppe_spm_term(tx6, samples, pol, k, n, m) +
                        ppe_spm_term(tx6, samples, pol, k, -m, -n) + ppe_spm_term(tx6, samples, pol, k, -n, -m);
                }
                F[row + rows * index++] = v;
            }
        }
        for (int m = -d_radius; m <= d_radius; ++m) {
            const int b = d_radius - (m < 0 ? -m : m);
            for (int n = -b; n <= b; ++n) F[row + rows * index++] = ppe_xpm_term(tx6, samples, pol, k, m, n);
        }
        F[row + rows * index] = tx6[(long long)(2 + pol) * samples + k];
    }
}

}  // namespace focss_hip
