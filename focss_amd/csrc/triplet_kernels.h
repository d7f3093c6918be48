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
//  k_gram        : G = F^H F over a column-major complex matrix, 32x32 tiles.
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

// G(i, j) = sum_r conj(F(r, i)) F(r, j); 32x32 output tile per workgroup,
// rows consumed in chunks of 32 through LDS
__global__ __launch_bounds__(256) void k_gram(const cx<double>* __restrict__ F, cx<double>* __restrict__ G,
                                              long long rows, int nf) {
    __shared__ cx<double> a[32][33];
    __shared__ cx<double> b[32][33];
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // ty in [0, 8)
    cx<double> acc[4];
    for (int q = 0; q < 4; ++q) acc[q] = mk<double>(0, 0);
    for (long long r0 = 0; r0 < rows; r0 += 32) {
        for (int q = 0; q < 4; ++q) {
            int c = ty + 8 * q;  // column within tile
            long long r = r0 + tx;
            a[c][tx] = (r < rows && i0 + c < nf) ? F[r + rows * (i0 + c)] : mk<double>(0, 0);
            b[c][tx] = (r < rows && j0 + c < nf) ? F[r + rows * (j0 + c)] : mk<double>(0, 0);
        }
        __syncthreads();
        for (int rr = 0; rr < 32; ++rr) {
            cx<double> bv = b[tx][rr];
            for (int q = 0; q < 4; ++q) acc[q] = acc[q] + conj(a[ty + 8 * q][rr]) * bv;
        }
        __syncthreads();
    }
    for (int q = 0; q < 4; ++q) {
        int i = i0 + ty + 8 * q, j = j0 + tx;
        if (i < nf && j < nf) G[(long long)i * nf + j] = acc[q];
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
