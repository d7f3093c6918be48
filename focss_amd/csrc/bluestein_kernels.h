// Fields whose length N is not a power of two (FFTW plans any N,
// src/focss/field.cpp:530-541).  Their DFTs go through Bluestein's chirp-z
// identity (Bluestein, 1970): with c_j = exp(-i pi j^2 / N),
//     X_k = sum_j x_j e^{-2 pi i jk/N} = c_k sum_j (x_j c_j) conj(c_{k-j}),
// a circular convolution of power-of-two length M >= 2N - 1, computed as
//     IDFT_M( DFT_M(a) . Bhat ),  Bhat = DFT_M(b) / M,  b_m = conj(c_|m|),
// with the library's power-of-two DFT kernels over rows of M values.  The
// inverse transform is conj(DFT(conj(X))).
//
// A link program on such a field runs as separate launches per operator:
//   time op        k_gen_time   (Kerr / span scale / ASE and the power
//                                reductions, the same time_op as the
//                                two-pass kernels, on the natural order)
//   spectral op    k_gen_in -> DFT_M -> k_gen_mul -> IDFT_M -> k_gen_mid
//                  -> DFT_M -> k_gen_mul -> IDFT_M -> k_gen_out
// k_gen_mid forms X = c . conv, plans the step (plan_spectral, as the
// spectral pass does), multiplies the dispersion/loss/gain factor and 1/N,
// and starts the inverse transform (conj, x c); k_gen_out ends it.  The
// field stays in the time domain between operators.
#pragma once
#include "ssfm_kernels.h"

namespace focss_hip {

// a waveform of a generic-length class
template <typename R>
struct GenWave {
    const cx<R>* chirp;      // c_j, j < N
    const cx<R>* bhat;       // Bhat_k, k < M
    const cx<R>* bhat_rows;  // Bhat in the row order of the fused kernels (M >= 2^12)
    // padded-convolution UNIFORM programs: Khat_s = DFT_M(K_s) / M in row
    // order for the step lengths kval[0], kval[1] (signed, as SpectralOp::s)
    const cx<R>* khat[2];
    double kval[2];
    long long row0;          // its first row of the work buffer (modes rows of M)
    int n;                   // samples per mode
    int pad;
};

template <typename R>
struct GenArgs {
    PassArgs p;               // field, control arrays, waves of this class, counter
    const GenWave<R>* gw;     // indexed by wave id
    cx<R>* work;              // [rows][M]
    int* active;              // per wave id: the spectral op of this counter runs
    int log_m;
    int modes;
    int chunks_per_row;       // M / chunk
};

constexpr int kGenChunk = 1024;  // values of one row per workgroup
constexpr int kGenTimeChunk = 512;  // samples per workgroup of k_gen_time

// workgroup -> (class wave k, mode, chunk)
template <typename R>
__device__ __forceinline__ void gen_locate(const GenArgs<R>& g, int& k, int& m, int& e0) {
    const int b = blockIdx.x;
    const int row = b / g.chunks_per_row;
    k = row / g.modes;
    m = row % g.modes;
    e0 = (b % g.chunks_per_row) * kGenChunk;
}

// a = x c (zero-padded to M) for waveforms whose program is still running
template <typename R>
__global__ __launch_bounds__(256) void k_gen_in(GenArgs<R> g) {
    int k, m, e0;
    gen_locate(g, k, m, e0);
    const int w_id = g.p.waves[k];
    const WaveState st = g.p.states[(g.p.counter & 1) * g.p.n_waves + w_id];
    const GenWave<R> gw = g.gw[w_id];
    const WaveStatic w = g.p.statics[w_id];
    const cx<R>* x = reinterpret_cast<const cx<R>*>(g.p.field) + w.offset + (long long)m * gw.n;
    cx<R>* a = g.work + (gw.row0 + m) * (1LL << g.log_m);
    const int M = 1 << g.log_m;
    // rows of finished programs are zeroed: the class's DFTs still run over them
    const int live = st.phase != PH_DONE;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < M; e += blockDim.x)
        a[e] = (live && e < gw.n) ? x[e] * gw.chirp[e] : mk<R>(0, 0);
}

// a *= Bhat
template <typename R>
__global__ __launch_bounds__(256) void k_gen_mul(GenArgs<R> g) {
    int k, m, e0;
    gen_locate(g, k, m, e0);
    const int w_id = g.p.waves[k];
    const GenWave<R> gw = g.gw[w_id];
    cx<R>* a = g.work + (gw.row0 + m) * (1LL << g.log_m);
    const int M = 1 << g.log_m;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < M; e += blockDim.x) a[e] = a[e] * gw.bhat[e];
}

// X = c conv; plan; Y = X D / N; start the inverse: a = conj(Y) c
template <typename R>
__global__ __launch_bounds__(256) void k_gen_mid(GenArgs<R> g) {
    int k, m, e0;
    gen_locate(g, k, m, e0);
    const int w_id = g.p.waves[k];
    const PassArgs& a = g.p;
    WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    const int was_done = st.phase == PH_DONE;
    const WaveStatic w = a.statics[w_id];
    const GenWave<R> gw = g.gw[w_id];
    const Reduction rd = a.red[(a.counter % 3) * a.n_waves + w_id];
    double sum = rd.sum;
    if (w.grid == GRID_LOGARITHMIC && st.phase == PH_SPAN_START) {
        // the time op's chunk sums in chunk order (run-to-run deterministic)
        const double* p = a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride;
        const int chunks = (gw.n + kGenTimeChunk - 1) / kGenTimeChunk;
        sum = 0;
        for (int t = 0; t < chunks; ++t) sum += p[t];
    } else if (w.energy_comp) {
        sum = comp_sum(st, rd, [&] {
            return tile_sum(a.partials + ((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride,
                            (gw.n + kGenTimeChunk - 1) / kGenTimeChunk);
        });
    }
    SpectralOp op = plan_spectral(w, st, bits_to_double(rd.peak_bits), sum);
    if (blockIdx.x % (g.chunks_per_row * g.modes) == 0 && threadIdx.x == 0) {
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
    const double arg = 0.5 * w.beta2 * op.s;
    const double mag = op.gain * op.loss / (double)gw.n;
    cx<R>* row = g.work + (gw.row0 + m) * (1LL << g.log_m);
    const int M = 1 << g.log_m;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < M; e += blockDim.x) {
        cx<R> v = mk<R>(0, 0);
        if (e < gw.n) {
            const cx<R> c = gw.chirp[e];
            const cx<R> y = (c * row[e]) * spectral_factor<R>(w, op, e, arg, mag);
            v = conj(y) * c;
        }
        row[e] = v;
    }
}

// x = conj(c conv) for the waveforms whose spectral op ran
template <typename R>
__global__ __launch_bounds__(256) void k_gen_out(GenArgs<R> g) {
    int k, m, e0;
    gen_locate(g, k, m, e0);
    const int w_id = g.p.waves[k];
    if (!g.active[w_id]) return;
    const GenWave<R> gw = g.gw[w_id];
    const WaveStatic w = g.p.statics[w_id];
    cx<R>* x = reinterpret_cast<cx<R>*>(g.p.field) + w.offset + (long long)m * gw.n;
    const cx<R>* row = g.work + (gw.row0 + m) * (1LL << g.log_m);
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < gw.n; e += blockDim.x) x[e] = conj(gw.chirp[e] * row[e]);
}

// the time-domain op on kGenTimeChunk samples (all modes) of one waveform,
// with the power reductions of the two-pass time pass
template <typename R, int M>
__global__ __launch_bounds__(256) void k_gen_time(GenArgs<R> g) {
    __shared__ double redbuf[2 * (256 / 64)];
    const PassArgs& a = g.p;
    const int chunks = a.tiles_per_wave;
    const int k = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    const int w_id = a.waves[k];
    const WaveState st = a.states[(a.counter & 1) * a.n_waves + w_id];
    if (st.top == TOP_NONE) return;
    const WaveStatic w = a.statics[w_id];
    const int n = g.gw[w_id].n;
    const int s0 = chunk * kGenTimeChunk;
    if (s0 >= n) return;  // whole workgroup: chunks beyond this waveform's length
    cx<R>* x = reinterpret_cast<cx<R>*>(a.field) + w.offset;
    double pk = 0, sm = 0;
    for (int s = s0 + threadIdx.x; s < s0 + kGenTimeChunk && s < n; s += 256) {
        cx<R> u[M];
#pragma unroll
        for (int mm = 0; mm < M; ++mm) u[mm] = x[(long long)mm * n + s];
        time_op<R, M>(u, w, st, s, pk, sm);
#pragma unroll
        for (int mm = 0; mm < M; ++mm) x[(long long)mm * n + s] = u[mm];
    }
    block_max_sum<256>(pk, sm, redbuf);
    if (threadIdx.x == 0) {
        Reduction* r = a.red + (a.counter % 3) * a.n_waves + w_id;
        atomicMax(&r->peak_bits, double_to_bits(pk));
        atomicAdd(&r->sum, sm);
        if (w.energy_comp && expected_energy(st) > 0) add_fixed_sum(r, sm, expected_energy(st));
        if (w.grid == GRID_LOGARITHMIC || (w.energy_comp && !(expected_energy(st) > 0)))
            a.partials[((long long)(a.counter % 3) * a.n_waves + w_id) * a.partial_stride + chunk] = sm;
    }
}

// ---- the standalone DFT of generic length (focss_fft) ----
// in: a = x c (forward) or conj(x) c (inverse), zero-padded; rows = modes
template <typename R>
__global__ __launch_bounds__(256) void k_blu_in(const cx<R>* __restrict__ x, cx<R>* __restrict__ work,
                                                const cx<R>* __restrict__ chirp, int n, int log_m, int inverse) {
    const int M = 1 << log_m;
    const int chunks = (M + kGenChunk - 1) / kGenChunk;
    const int row = blockIdx.x / chunks, e0 = (blockIdx.x % chunks) * kGenChunk;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < M; e += blockDim.x) {
        cx<R> v = mk<R>(0, 0);
        if (e < n) {
            const cx<R> xv = x[(long long)row * n + e];
            v = (inverse ? conj(xv) : xv) * chirp[e];
        }
        work[(long long)row * M + e] = v;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_blu_mul(cx<R>* __restrict__ work, const cx<R>* __restrict__ bhat,
                                                 int log_m) {
    const int M = 1 << log_m;
    const int chunks = (M + kGenChunk - 1) / kGenChunk;
    const int row = blockIdx.x / chunks, e0 = (blockIdx.x % chunks) * kGenChunk;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < M; e += blockDim.x)
        work[(long long)row * M + e] = work[(long long)row * M + e] * bhat[e];
}

// out: X = c conv x scale (forward) or conj(c conv) (inverse)
template <typename R>
__global__ __launch_bounds__(256) void k_blu_out(const cx<R>* __restrict__ work, cx<R>* __restrict__ x,
                                                 const cx<R>* __restrict__ chirp, int n, int log_m, int inverse,
                                                 double scale) {
    const int M = 1 << log_m;
    const int chunks = (M + kGenChunk - 1) / kGenChunk;
    const int row = blockIdx.x / chunks, e0 = (blockIdx.x % chunks) * kGenChunk;
    const R sc = (R)scale;
    for (int e = e0 + threadIdx.x; e < e0 + kGenChunk && e < n; e += blockDim.x) {
        const cx<R> v = chirp[e] * work[(long long)row * M + e];
        x[(long long)row * n + e] = inverse ? conj(v) : v * sc;
    }
}

}  // namespace focss_hip
