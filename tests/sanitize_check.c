/* Memory and undefined-behaviour check of the CPU oracle: every entry point
 * of oracle/focss_oracle.h that the tests use, on small seeded inputs, with
 * the edge cases the tests rely on (lengths that are powers of two, small
 * primes, a Bluestein length, one and two modes, the three SSFM grids, ASE,
 * the triplet equalizers' training, learned DBP and the triplet MLP).
 * tests/test_sanitize_cpu.py builds it with the oracle under
 * -fsanitize=address,undefined and requires a clean exit; values are
 * checked elsewhere. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "focss_oracle.h"

static unsigned long long state = 0x9E3779B97F4A7C15ull;

static double uniform(void) {
    state = state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(state >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

static orc_c128* field(int n, double scale) {
    orc_c128* f = (orc_c128*)malloc(sizeof(orc_c128) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        f[i].re = scale * uniform();
        f[i].im = scale * uniform();
    }
    return f;
}

static double sink = 0;  /* results are consumed so that nothing is optimised away */

static void use(const orc_c128* f, int n) {
    for (int i = 0; i < n; ++i) sink += f[i].re - f[i].im;
}

static void check_fft(void) {
    const int lengths[] = {1, 2, 3, 5, 7, 12, 64, 97, 100, 1000, 1 << 12, 4099, 3 << 10};
    for (unsigned t = 0; t < sizeof(lengths) / sizeof(lengths[0]); ++t)
        for (int modes = 1; modes <= 2; ++modes) {
            const int n = lengths[t];
            orc_c128* x = field(modes * n, 1.0);
            orc_fft_inplace(modes, n, x);
            orc_ifft_inplace(modes, n, x);
            if (n <= 128) {
                orc_c128* y = field(n, 0.0);
                orc_dft_naive(-1, n, x, y);
                use(y, n);
                free(y);
            }
            use(x, modes * n);
            free(x);
        }
}

static void check_ssfm(void) {
    const orc_fiber fib = {4.6e-5, -2.17e-26, 1.3e-3, 2e3, 1.55e-6, 1.0};
    const orc_grid grids[3] = {{ORC_UNIFORM, 3, 0, 0}, {ORC_LOGARITHMIC, 0, 1e-2, 1e3}, {ORC_ADAPTIVE, 0, 1e-2, 1e3}};
    const int lengths[] = {64, 1000, 1 << 12};
    for (int g = 0; g < 3; ++g)
        for (int t = 0; t < 3; ++t)
            for (int modes = 1; modes <= 2; ++modes) {
                const int n = lengths[t];
                orc_c128* x = field(modes * n, 0.05);
                int steps = 0;
                orc_ssfm_run(&fib, &grids[g], g == 1 ? -1.0 : 1.0, modes, n, 1.0 / 64e9, x, &steps);
                orc_amplify(39.8, 3.16, 193.4e12, 1.0 / 64e9, 7, 1, modes, n, x);
                orc_drop_power(39.8, modes, n, x);
                sink += orc_peak_power(modes, n, x) + orc_average_power(modes, n, x) + steps;
                free(x);
            }
    const int n = 2048;
    orc_c128* x = field(n, 0.05);
    int per_span[2] = {0, 0};
    orc_forward_propagation(&fib, 2, 1e-2, 1e3, 39.8, 3.16, 193.4e12, 11, 1, n, 1.0 / 64e9, x, per_span);
    orc_backward_propagation(&fib, 2, 4, 1e-2, 1e3, 39.8, 1, n, 1.0 / 64e9, x);
    orc_backward_propagation(&fib, 2, 0, 1e-2, 1e3, 39.8, 1, n, 1.0 / 64e9, x);
    orc_cd_compensation(&fib, 2, 1, n, 1.0 / 64e9, x);
    use(x, n);
    free(x);
}

static void check_dsp_and_scoring(void) {
    for (int modes = 1; modes <= 2; ++modes) {
        const int n_sym = 256, osf = 4, n = n_sym * osf;
        orc_c128* sym = field(modes * n_sym, 1.0);
        orc_c128* sig = field(modes * n, 0.0);
        orc_c128* back = field(modes * n_sym, 0.0);
        orc_c128* dec = field(modes * n / 2, 0.0);
        orc_c128* sel = field(modes * n, 0.0);
        orc_c128* pulse = field(n, 0.0);
        orc_rrc_pulse(n_sym, osf, 0.1, pulse);
        orc_rrc_shaping(modes, n_sym, sym, osf, 0.1, 1.0, sig);
        orc_rrc_sampling(modes, n, sig, osf, 0.1, 1.0, back);
        orc_decimate(modes, n, sig, 2, dec);
        orc_frequency_shift(modes, n, 1.0 / 128e9, 5e9, sig);
        orc_bandpass_filter(modes, n, 1.0 / 128e9, 20e9, sig);
        orc_wdm_select(modes, n, 1.0 / 128e9, sig, 1, 50e9, sel);
        orc_filter_inplace(modes, n, sig, n_sym, pulse);
        int* codes = (int*)malloc(sizeof(int) * (size_t)(modes * n_sym));
        orc_hd_sequence_16qam(modes, n_sym, back, codes);
        sink += codes[0] + orc_demodulate_16qam(0.1, -0.3);
        sink += orc_ber_16qam(modes, n_sym, sym, back, 0) + orc_ber_16qam(modes, n_sym, sym, back, 8);
        sink += orc_q2_factor(modes, n_sym, sym, back, 0) + orc_q2_factor(modes, n_sym, sym, back, 8);
        sink += orc_pse_estimate(modes * n_sym, sym, back, 360) + orc_homothety_estimate(modes * n_sym, sym, back);
        orc_c128 s = orc_scalar_estimate(modes * n_sym, sym, back);
        sink += s.re;
        use(sel, modes * n);
        use(dec, modes * n / 2);
        free(codes);
        free(sym);
        free(sig);
        free(back);
        free(dec);
        free(sel);
        free(pulse);
    }
}

static void check_triplets(void) {
    for (int modes = 1; modes <= 2; ++modes)
        for (int type = 0; type <= 2; ++type) {
            const int r = 3, n = 512, nw = (2 * r + 1) * (2 * r + 1);
            orc_c128* tx = field(modes * n, 1.0);
            orc_c128* rx = field(modes * n, 1.0);
            orc_c128* w = field(nw, 0.0);
            orc_c128* out = field(modes * n, 0.0);
            orc_c128 ic, lw, zero = {0, 0}, term;
            orc_sino_train(r, type, zero, modes, n, tx, rx, w, &ic, &lw);
            orc_sino_cubic(r, w, modes, n, rx, out);
            orc_sino_term(modes, n, rx, modes - 1, n / 2, 1, -2, &term);
            const int nf = orc_sino_n_features(r, type), rows = modes * (n - 2 * r);
            orc_c128* feat = field(rows * nf, 0.0);
            orc_c128* resp = field(rows, 0.0);
            orc_sino_features(r, type, zero, modes, n, tx, rx, feat, resp);
            use(out, modes * n);
            use(resp, rows);
            sink += term.re + ic.re + lw.re;
            free(tx);
            free(rx);
            free(w);
            free(out);
            free(feat);
            free(resp);
        }
    const int c = 2, d = 1, n = 256;
    orc_c128* tx6 = field(6 * n, 1.0);
    orc_c128* rx6 = field(6 * n, 1.0);
    orc_c128* C = field((2 * c + 1) * (2 * c + 1), 0.0);
    orc_c128* D = field((2 * d + 1) * (2 * d + 1), 0.0);
    orc_c128* out2 = field(2 * n, 0.0);
    orc_c128 ic, lw;
    orc_wdmppe_train(c, d, n, tx6, rx6, C, D, &ic, &lw);
    orc_wdmppe_spm(c, C, n, rx6, out2);
    orc_wdmppe_xpm(d, D, n, rx6, out2);
    use(out2, 2 * n);
    const int rows = 12, cols = 5;
    orc_c128* a = field(rows * cols, 1.0);
    orc_c128* b = field(rows, 1.0);
    orc_c128* xs = field(cols, 0.0);
    orc_complex_lstsq(rows, cols, a, b, xs);
    use(xs, cols);
    free(tx6);
    free(rx6);
    free(C);
    free(D);
    free(out2);
    free(a);
    free(b);
    free(xs);
}

static void check_learned(void) {
    for (int modes = 1; modes <= 2; ++modes)
        for (int ktaps = 1; ktaps <= 5; ktaps += 4) {
            const int steps = 3, taps = 7, n = 256;
            orc_c128* fir = field(steps * taps, 0.2);
            double* kerr = (double*)malloc(sizeof(double) * steps * ktaps);
            for (int i = 0; i < steps * ktaps; ++i) kerr[i] = -0.01 * (1 + uniform());
            orc_c128* x = field(modes * n, 1.0);
            orc_c128* y = field(modes * n, 0.0);
            orc_c128* tgt = field(modes * n, 1.0);
            orc_c128* gf = field(steps * taps, 0.0);
            double* gk = (double*)calloc((size_t)(steps * ktaps), sizeof(double));
            sink += orc_ldbp_forward(steps, taps, ktaps, modes, fir, kerr, n, x, y, tgt);
            sink += orc_ldbp_backward(steps, taps, ktaps, modes, fir, kerr, n, x, tgt, gf, gk);
            use(gf, steps * taps);
            free(fir);
            free(kerr);
            free(x);
            free(y);
            free(tgt);
            free(gf);
            free(gk);
        }
    const int np = orc_tmlp_n_params(), modes = 2, symbols = 64, batch = 2;
    double* p = (double*)malloc(sizeof(double) * (size_t)np);
    double* g = (double*)calloc((size_t)np, sizeof(double));
    for (int i = 0; i < np; ++i) p[i] = 0.01 * uniform();
    orc_c128* x = field(batch * modes * symbols, 1.0);
    orc_c128* y = field(batch * modes * symbols, 0.0);
    orc_c128* tgt = field(batch * modes * symbols, 1.0);
    for (int bf16 = 0; bf16 <= 1; ++bf16) {
        sink += orc_tmlp_forward(modes, symbols, batch, p, bf16, x, y, tgt);
        sink += orc_tmlp_backward(modes, symbols, batch, p, bf16, x, tgt, g);
    }
    use(y, batch * modes * symbols);
    free(p);
    free(g);
    free(x);
    free(y);
    free(tgt);
}

int main(void) {
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
    double z0, z1;
    orc_philox4x32_10(ctr, key, out);
    orc_gauss_pair(42, 3, 1000, &z0, &z1);
    sink += out[0] + z0 + z1;
    check_fft();
    check_ssfm();
    check_dsp_and_scoring();
    check_triplets();
    check_learned();
    printf("sanitize check done (%d)\n", isfinite(sink) ? 1 : 0);
    return 0;
}
