// Learned few-step digital backpropagation (no counterpart in the
// reference; judged against plain DBP).  Step j of J:
//     y(t) = sum_k h_{j,k} x((t - k + c) mod N),  c = K / 2      (complex FIR)
//     z(t) = y(t) exp(i theta_j |y(t)|^2)                        (Kerr phase)
// Loss = sum over the batch of (1/(N B)) sum_t |z_J(t) - target(t)|^2.
// Gradients are Wirtinger dL/d(conj h) and dL/dtheta:
//     q = -Im(conj(G_z) z),  G_y = G_z e^{-i theta p} + 2 theta q y,
//     dL/dtheta += 2 p q,    G_h_k += G_y(t) conj(x(t - k + c)),
//     G_x(s) = sum_k conj(h_k) G_y(s + k - c).
// One 1024-thread workgroup owns one waveform (N <= 16384 complex64 in LDS)
// and runs every step on its CU; the forward pass can store x_j and y_j for
// the backward pass.  A step's outputs depend on neighbouring inputs, so they
// are staged in a per-waveform global scratch (L2-resident) and copied back
// after a barrier -- a second LDS buffer would not fit beside the first.
#pragma once
#include "cx.h"

namespace focss_hip {

constexpr int LDBP_NT = 1024;
constexpr int LDBP_MAX_N = 16384;

struct LdbpArgs {
    const cx<float>* in;      // [B][N]
    cx<float>* out;           // [B][N] (may be null)
    const cx<float>* target;  // [B][N] (may be null)
    const cx<float>* fir;     // [J][K]
    const float* theta;       // [J]
    cx<float>* xs;            // [B][J][N] FIR inputs (training) or null
    cx<float>* ys;            // [B][J][N] FIR outputs (training) or null
    double* loss;             // [1] accumulated, may be null
    double* grad;             // [J*K*2 + J] accumulated: Re/Im of G_h then dtheta
    cx<float>* zs;            // [B][N] scratch: step outputs staged before they overwrite LDS
    int J, K, N, B;
};

__device__ __forceinline__ double ldbp_block_sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < LDBP_NT / 64; ++i) s += red[i];
    return s;  // valid in thread 0
}

template <int K>
__global__ __launch_bounds__(LDBP_NT) void k_ldbp_fwd(LdbpArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cx<float>* buf = reinterpret_cast<cx<float>*>(smem_raw);
    __shared__ double red[LDBP_NT / 64];
    __shared__ cx<float> h[K];
    const int N = a.N, mask = N - 1, c = K / 2;
    const int b = blockIdx.x;
    const cx<float>* in = a.in + (long long)b * N;
    for (int t = threadIdx.x; t < N; t += LDBP_NT) buf[t] = in[t];
    __syncthreads();
    cx<float>* zs = a.zs + (long long)b * N;
    for (int j = 0; j < a.J; ++j) {
        if (threadIdx.x < K) h[threadIdx.x] = a.fir[j * K + threadIdx.x];
        __syncthreads();
        const float th = a.theta[j];
        for (int t = threadIdx.x; t < N; t += LDBP_NT) {
            cx<float> y = mk<float>(0, 0);
#pragma unroll 16
            for (int k = 0; k < K; ++k) y = y + h[k] * buf[(t - k + c) & mask];
            if (a.ys) {
                a.xs[((long long)b * a.J + j) * N + t] = buf[t];
                a.ys[((long long)b * a.J + j) * N + t] = y;
            }
            float s, co;
            sincosf(th * norm2(y), &s, &co);
            zs[t] = y * mk<float>(co, s);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < N; t += LDBP_NT) buf[t] = zs[t];
        __syncthreads();
    }
    double part = 0;
    for (int t = threadIdx.x; t < N; t += LDBP_NT) {
        if (a.out) a.out[(long long)b * N + t] = buf[t];
        if (a.target) {
            cx<float> e = buf[t] - a.target[(long long)b * N + t];
            part += (double)norm2(e);
        }
    }
    if (a.loss) {
        double s = ldbp_block_sum(part, red);
        if (threadIdx.x == 0) atomicAdd(a.loss, s / ((double)N * a.B));
    }
}

// needs the forward pass run with xs / ys stored; out holds z_J
template <int K>
__global__ __launch_bounds__(LDBP_NT) void k_ldbp_bwd(LdbpArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cx<float>* g = reinterpret_cast<cx<float>*>(smem_raw);
    __shared__ double red[LDBP_NT / 64];
    __shared__ cx<float> h[K];
    constexpr int KC = 8;  // taps per gradient sweep (register budget)
    const int N = a.N, mask = N - 1, c = K / 2;
    const int b = blockIdx.x;
    const float scale = 1.0f / ((float)N * (float)a.B);
    cx<float>* zs = a.zs + (long long)b * N;
    for (int t = threadIdx.x; t < N; t += LDBP_NT)
        g[t] = (a.out[(long long)b * N + t] - a.target[(long long)b * N + t]) * scale;
    __syncthreads();
    for (int j = a.J - 1; j >= 0; --j) {
        const cx<float>* x = a.xs + ((long long)b * a.J + j) * N;
        const cx<float>* y = a.ys + ((long long)b * a.J + j) * N;
        const float th = a.theta[j];
        // Kerr backward, in place (each thread owns its samples)
        double dth = 0;
        for (int t = threadIdx.x; t < N; t += LDBP_NT) {
            cx<float> yv = y[t];
            float p = norm2(yv);
            float s, co;
            sincosf(th * p, &s, &co);
            cx<float> zv = yv * mk<float>(co, s);
            cx<float> gz = g[t];
            float q = -(gz.x * zv.y - gz.y * zv.x);  // -Im(conj(gz) z)
            dth += 2.0 * (double)p * (double)q;
            g[t] = gz * mk<float>(co, -s) + yv * (2.0f * th * q);
        }
        __syncthreads();
        // FIR tap gradients G_h_k = sum_t G_y(t) conj(x(t - k + c)), KC taps per
        // sweep; per-thread partial sums in float, block sums in double
        double* gr = a.grad + (long long)j * K * 2;
        for (int k0 = 0; k0 < K; k0 += KC) {
            float dre[KC], dim[KC];
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                dre[q] = 0;
                dim[q] = 0;
            }
            for (int t = threadIdx.x; t < N; t += LDBP_NT) {
                cx<float> gy = g[t];
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    if (k0 + q < K) {
                        cx<float> v = gy * conj(x[(t - (k0 + q) + c) & mask]);
                        dre[q] += v.x;
                        dim[q] += v.y;
                    }
                }
            }
#pragma unroll 1
            for (int q = 0; q < KC && k0 + q < K; ++q) {
                double sr = ldbp_block_sum((double)dre[q], red);
                if (threadIdx.x == 0) atomicAdd(&gr[2 * (k0 + q)], sr);
                double si = ldbp_block_sum((double)dim[q], red);
                if (threadIdx.x == 0) atomicAdd(&gr[2 * (k0 + q) + 1], si);
            }
        }
        double st = ldbp_block_sum(dth, red);
        if (threadIdx.x == 0) atomicAdd(&a.grad[(long long)a.J * K * 2 + j], st);
        // input gradient G_x(s) = sum_k conj(h_k) G_y(s + k - c), staged in zs
        if (threadIdx.x < K) h[threadIdx.x] = conj(a.fir[j * K + threadIdx.x]);
        __syncthreads();
        for (int s = threadIdx.x; s < N; s += LDBP_NT) {
            cx<float> acc = mk<float>(0, 0);
#pragma unroll 16
            for (int k = 0; k < K; ++k) acc = acc + h[k] * g[(s + k - c) & mask];
            zs[s] = acc;
        }
        __syncthreads();
        for (int s = threadIdx.x; s < N; s += LDBP_NT) g[s] = zs[s];
        __syncthreads();
    }
}

// Adam (Kingma & Ba, 2015) on float parameters with double gradients;
// p holds the FIR taps (re, im interleaved) followed by theta.
__global__ void k_adam(float* __restrict__ p, const double* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float beta1, float beta2, float eps, int t,
                       const float* __restrict__ lr_scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the taps carry the Wirtinger gradient dL/d(conj h) = (dL/dRe + i dL/dIm) / 2;
    // Adam normalises every coordinate, so the factor 2 against theta's real
    // gradient does not change the update
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
