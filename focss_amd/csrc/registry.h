// Kernel registries of the SSFM passes and the standalone DFT: tables of
// kernel pointers indexed by (log2 N, modes, tile variant), filled from
// several translation units (registry_*.hip) so the ~300 instantiations
// compile in parallel.
#pragma once
#include "ssfm_kernels.h"
#include "fft_kernels.h"
#include "bluestein_fused.h"

namespace focss_hip {

template <typename R>
using KernelFn = void (*)(PassArgs, Tables<R>);

// variants 0-2: one tile per workgroup; 3-5: the same tiles, two per
// workgroup (the second one's loads in flight during the first's DFTs)
constexpr int kTileVariants = 6;

#ifndef FOCSS_SPLIT_SHIFT
#define FOCSS_SPLIT_SHIFT 0
#endif
// log2 N1 of the two-pass split N = N1 x N2 of the SSFM passes: the time
// pass runs length-N1 DFTs down columns, the spectral pass length-N2 DFTs
// along rows.
constexpr int pass_log_n1(int log_n) { return log_n / 2 - (log_n >= 16 ? FOCSS_SPLIT_SHIFT : 0); }
constexpr int kTileElems[kTileVariants] = {4096, 2048, 1024, 4096, 2048, 1024};
constexpr int kTilesPerGroup[kTileVariants] = {1, 1, 1, 2, 2, 2};

template <typename R>
struct Registry {
    // resident[log_n][modes-1]
    KernelFn<R> resident[21][2] = {};
    // time[log_n][modes-1][variant], spec[log_n][variant]
    KernelFn<R> time_k[21][2][kTileVariants] = {};
    KernelFn<R> spec_k[21][kTileVariants] = {};
    int time_cols[21][2][kTileVariants] = {};
    int spec_rows[21][kTileVariants] = {};
    int time_threads[21][2][kTileVariants] = {};
    int spec_threads[21][kTileVariants] = {};
    int time_tpb[21][2][kTileVariants] = {};  // tiles per workgroup
    int spec_tpb[21][kTileVariants] = {};
};

template <typename R, int N, int M>
void reg_resident(Registry<R>& r) {
    r.resident[ilog2<N>::value][M - 1] = &k_resident<R, N, M>;
}

template <typename R, int LOGN, int M, int V>
void reg_time(Registry<R>& r) {
    constexpr int N1 = 1 << pass_log_n1(LOGN);
    constexpr int N2 = 1 << (LOGN - pass_log_n1(LOGN));
    constexpr int C0 = kTileElems[V] / (N1 * M);
    constexpr int C = C0 < 1 ? 0 : (C0 > N2 ? N2 : C0);
    constexpr int T = kTilesPerGroup[V];
    if constexpr (C >= 1 && (N2 / (C < 1 ? 1 : C)) % T == 0) {
        r.time_k[LOGN][M - 1][V] = &k_time_pass<R, N1, C, M, T>;
        r.time_cols[LOGN][M - 1][V] = C;
        r.time_threads[LOGN][M - 1][V] = pass_threads(C * N1 / 8);
        r.time_tpb[LOGN][M - 1][V] = T;
    }
}

template <typename R, int LOGN, int V>
void reg_spec(Registry<R>& r) {
    constexpr int N2 = 1 << (LOGN - pass_log_n1(LOGN));
    constexpr int N1 = 1 << pass_log_n1(LOGN);
    constexpr int S0 = kTileElems[V] / N2;
    constexpr int S = S0 < 1 ? 0 : (S0 > N1 ? N1 : S0);
    constexpr int T = kTilesPerGroup[V];
    // tiles per waveform are modes N1 / S: a multiple of T for either mode count
    if constexpr (S >= 1 && (N1 / (S < 1 ? 1 : S)) % T == 0) {
        r.spec_k[LOGN][V] = &k_spec_pass<R, N2, S, T>;
        r.spec_rows[LOGN][V] = S;
        r.spec_threads[LOGN][V] = pass_threads(S * N2 / 8);
        r.spec_tpb[LOGN][V] = T;
    }
}

template <typename R, int LOGN>
void reg_two_pass(Registry<R>& r) {
    reg_time<R, LOGN, 1, 0>(r);
    reg_time<R, LOGN, 1, 1>(r);
    reg_time<R, LOGN, 1, 2>(r);
    reg_time<R, LOGN, 1, 3>(r);
    reg_time<R, LOGN, 1, 4>(r);
    reg_time<R, LOGN, 1, 5>(r);
    reg_time<R, LOGN, 2, 0>(r);
    reg_time<R, LOGN, 2, 1>(r);
    reg_time<R, LOGN, 2, 2>(r);
    reg_time<R, LOGN, 2, 3>(r);
    reg_time<R, LOGN, 2, 4>(r);
    reg_time<R, LOGN, 2, 5>(r);
    reg_spec<R, LOGN, 0>(r);
    reg_spec<R, LOGN, 1>(r);
    reg_spec<R, LOGN, 2>(r);
    reg_spec<R, LOGN, 3>(r);
    reg_spec<R, LOGN, 4>(r);
    reg_spec<R, LOGN, 5>(r);
}

// standalone DFT kernels: [log_n][sign 0 = forward, 1 = inverse]
template <typename R>
struct FftRegistry {
    void (*whole[21][2])(cx<R>*, double, Tables<R>) = {};
    void (*cols[21][2])(const cx<R>*, cx<R>*, int, int, Tables<R>) = {};
    void (*rows[21][2])(const cx<R>*, cx<R>*, int, int, double, Tables<R>) = {};
    int ncols[21] = {}, nrows[21] = {};
};

template <typename R, int LOGN>
void reg_fft_whole(FftRegistry<R>& r) {
    r.whole[LOGN][0] = &k_fft_whole<R, (1 << LOGN), -1>;
    r.whole[LOGN][1] = &k_fft_whole<R, (1 << LOGN), +1>;
}

template <typename R, int LOGN>
void reg_fft_two(FftRegistry<R>& r) {
    constexpr int N1 = 1 << (LOGN / 2);
    constexpr int N2 = 1 << (LOGN - LOGN / 2);
    constexpr int TILE = sizeof(R) == 8 ? 2048 : 4096;
    constexpr int C = TILE / N1 > N2 ? N2 : TILE / N1;
    constexpr int S = TILE / N2;
    r.cols[LOGN][0] = &k_fft_cols<R, N1, C, -1>;
    r.cols[LOGN][1] = &k_fft_cols<R, N1, C, +1>;
    r.rows[LOGN][0] = &k_fft_rows<R, N2, S, -1>;
    r.rows[LOGN][1] = &k_fft_rows<R, N2, S, +1>;
    r.ncols[LOGN] = C;
    r.nrows[LOGN] = S;
}

// fused Bluestein link kernels: [log_m], [log_m][modes-1][kind]
template <typename R>
struct BluRegistry {
    void (*rows[21])(BluArgs<R>, int, Tables<R>) = {};
    void (*rows_pad[21])(BluArgs<R>, int, Tables<R>) = {};  // padded-convolution spectral op
    void (*cols[21][2][5])(BluArgs<R>, Tables<R>) = {};
    int row_tiles[21] = {};     // row-kernel workgroups per mode (N1 / S)
    int col_tiles[21][2] = {};  // column-kernel workgroups per waveform (N2 / C)
    int row_threads[21] = {};   // workgroup sizes the kernels are built for
    int col_threads[21][2] = {};
};

template <typename R, int LOGM, int MODES>
void reg_blu_modes(BluRegistry<R>& r) {
    using Sh = BluShape<R, LOGM>;
    r.cols[LOGM][MODES - 1][BLU_ENTRY] = &k_blu_cols<R, LOGM, MODES, BLU_ENTRY>;
    r.cols[LOGM][MODES - 1][BLU_MID] = &k_blu_cols<R, LOGM, MODES, BLU_MID>;
    r.cols[LOGM][MODES - 1][BLU_OUT] = &k_blu_cols<R, LOGM, MODES, BLU_OUT>;
    r.cols[LOGM][MODES - 1][BLU_PAD_ENTRY] = &k_blu_cols<R, LOGM, MODES, BLU_PAD_ENTRY>;
    r.cols[LOGM][MODES - 1][BLU_PAD_OUT] = &k_blu_cols<R, LOGM, MODES, BLU_PAD_OUT>;
    r.col_tiles[LOGM][MODES - 1] = Sh::N2 / Sh::template C<MODES>;
    r.col_threads[LOGM][MODES - 1] = pass_threads(Sh::template C<MODES> * Sh::N1 / 8);
}

template <typename R, int LOGM>
void reg_blu(BluRegistry<R>& r) {
    using Sh = BluShape<R, LOGM>;
    static_assert(Sh::N1 % Sh::S == 0 && Sh::N2 % Sh::template C<1> == 0 && Sh::N2 % Sh::template C<2> == 0,
                  "Bluestein tile shape");
    r.rows[LOGM] = &k_blu_rows<R, LOGM, false>;
    r.rows_pad[LOGM] = &k_blu_rows<R, LOGM, true>;
    r.row_tiles[LOGM] = Sh::N1 / Sh::S;
    r.row_threads[LOGM] = pass_threads(Sh::S * Sh::N2 / 8);
    reg_blu_modes<R, LOGM, 1>(r);
    reg_blu_modes<R, LOGM, 2>(r);
}

// the parts, one per translation unit
void fill_resident(Registry<double>& r);
void fill_resident(Registry<float>& r);
void fill_two_pass_small(Registry<double>& r);  // 2^12 .. 2^15
void fill_two_pass_small(Registry<float>& r);
void fill_two_pass_large(Registry<double>& r);  // 2^16 .. 2^20
void fill_two_pass_large(Registry<float>& r);
void fill_fft(FftRegistry<double>& r);
void fill_fft(FftRegistry<float>& r);
void fill_blu(BluRegistry<double>& r);
void fill_blu(BluRegistry<float>& r);

template <typename R>
const Registry<R>& reg() {
    static const Registry<R> r = [] {
        Registry<R> x;
        fill_resident(x);
        fill_two_pass_small(x);
        fill_two_pass_large(x);
        return x;
    }();
    return r;
}

template <typename R>
const FftRegistry<R>& fft_reg() {
    static const FftRegistry<R> r = [] {
        FftRegistry<R> x;
        fill_fft(x);
        return x;
    }();
    return r;
}

template <typename R>
const BluRegistry<R>& blu_reg() {
    static const BluRegistry<R> r = [] {
        BluRegistry<R> x;
        fill_blu(x);
        return x;
    }();
    return r;
}

}  // namespace focss_hip
