// Memory and undefined-behaviour check of the C++ mirror's host-side code
// (no call here reaches the GPU): Field construction, copies, moves,
// views and host arithmetic; the scoring and decision functions; the RRC
// pulse; the transmission text format and the save_field record format
// (round trips through memory); Fiber arithmetic; Vector views.
// tests/test_sanitize_cpu.py compiles it together with the host sources
// under -fsanitize=address,undefined and requires a clean exit.
#include <cmath>
#include <cstdio>
#include <sstream>

#include "focss/field.h"
#include "focss/functions.h"
#include "focss/module/fiber.h"
#include "focss/processing/modulation.h"
#include "focss/vector.h"
#include "tasks/field_io.h"

using namespace focss;

namespace {
unsigned long long g_state = 0x243F6A8885A308D3ull;
double uniform() {
    g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(g_state >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}
Field random_field(int modes, int samples) {
    Field f(Domain{modes, samples, 1.0 / 64e9, 193.4e12});
    for (int i = 0; i < f.size(); ++i) f[i] = complex_t(uniform(), uniform());
    return f;
}
double sink = 0;
}  // namespace

int main() {
    for (int modes = 1; modes <= 2; ++modes) {
        const int n = 512;
        Field tx = random_field(modes, n), rx = random_field(modes, n);
        Field copy(tx), moved(std::move(copy)), assigned;
        assigned = moved;
        assigned *= 0.5;
        assigned *= complex_t(0.0, 1.0);
        assigned += rx;
        assigned -= tx;
        assigned /= 2.0;
        RealVector gains(n, 1.5);  // one factor per sample, all modes
        assigned *= gains;
        ComplexVector mode0 = assigned(0);
        sink += mode0[0].real() + assigned.power(3) + assigned.peak_power() + assigned.average_power();
        assigned.peak_normalize(2.0).average_normalize(1.0);
        sink += assigned.temporal_power()[1] + assigned.time_grid()[2] + assigned.frequency_grid()[3];
        Field up = assigned.upsample(2), down = up.downsample(2);
        down.fft_shift();
        sink += up[1].real() + down[2].imag();

        sink += q2_factor(tx, rx) + q2_factor(tx, rx, 8) + evm2_factor(tx, rx) + evm2_factor(tx, rx, 8);
        sink += bit_error_rate_16qam(tx, rx) + bit_error_rate_16qam(tx, rx, 8) + symbol_error_rate_16qam(tx, rx, 8);
        Vector<int> codes = hd_sequence_16qam(rx);
        sink += codes[0] + demodulate_16qam(complex_t(0.0, 2.0 / std::sqrt(10.0)));

        std::stringstream text;
        save_transmission(text, tx, rx);
        Field tx2(tx.domain()), rx2(rx.domain());  // filled in place, as the reference's reader
        load_transmission(text, &tx2, &rx2);
        sink += tx2[n - 1].real() + rx2[0].imag();

        Vector<Field> envelopes(3);
        envelopes[0] = tx;
        envelopes[1] = rx;
        envelopes[2] = tx;
        std::stringstream records;
        save_field(records, envelopes);
        FieldFileHeader h;
        Field back = load_field(records, &h, 1.0 / 64e9);
        sink += back[back.size() - 1].real() + h.channels;
    }
    ComplexVector pulse = pulse::rrc(64, 4, 0.1);
    sink += pulse[128].real();
    Fiber f{4.6e-5, -2.17e-26, 1.3e-3, 80e3, 1.55e-6, 1.0};
    Fiber g = (-f) * 0.5, h = f * -2.0, l = f.linear(), d = f.dispersive(), nl = f.nonlinear();
    sink += g.length + h.alpha + l.gamma + d.beta2 + nl.gamma;
    sink += db_to_linear(3.0) + dbm_to_watts(0.0) + db_to_natural(0.2) + sinc(0.3) + i_exp(0.7).imag();
    ComplexVector store(16, complex_t(1.0, 0.0));
    ComplexVector view = ComplexVector::proxy(8, store.raw() + 4);
    view *= complex_t(2.0, 0.0);
    sink += store[5].real();
    std::printf("host sanitize check done (%d)\n", std::isfinite(sink) ? 1 : 0);
    return 0;
}
