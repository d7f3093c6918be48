// Transmitter / receiver helpers (reference interface
// src/focss/processing/modulation.h:8-45; behaviour of modulation.cpp:9-277).
#include "focss/processing/modulation.h"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>

#include "focss/runtime.h"

namespace focss {

namespace {

// Gray 16QAM (modulation.cpp:9-17): code = 4 q + i, each axis Gray-mapped
// 0 -> -3, 1 -> -1, 3 -> +1, 2 -> +3, scaled to unit mean power.
const double kAxisLevel[4] = {-3.0, -1.0, 3.0, 1.0};

// The device scratch a Field visits for one C-ABI call: uploaded on
// construction, downloaded on request.
struct Staged {
    DeviceBuffer buf;
    explicit Staged(const Field& f) : buf(f.raw(), sizeof(complex_t) * (size_t)f.size()) {}
    explicit Staged(size_t values) : buf(sizeof(complex_t) * values) {}
    Field fetch(const Domain& d) const {
        Field out(d);
        buf.download(out.raw());
        return out;
    }
};

// average power per mode, the decision scale of the 16QAM helpers
double power_per_mode(const Field& f) { return f.average_power() / f.modes(); }

unsigned bit_distance(unsigned a, unsigned b) { return (unsigned)__builtin_popcount(a ^ b); }

}  // namespace

complex_t gray_16qam(const unsigned int& code) {
    // divided (not multiplied by 1/sqrt(10)) so that the points, and with
    // them the decisions on boundaries, round as the reference's table does
    const double root10 = std::sqrt(10.0);
    return complex_t(kAxisLevel[code & 3u] / root10, kAxisLevel[(code >> 2) & 3u] / root10);
}

// pulse::sech (modulation.cpp:19-31): 1/cosh(t / width) / width on a grid
// of n points spanning t = -10 .. +10
ComplexVector pulse::sech(const int& n_samples, const double& width) {
    ComplexVector out(n_samples);
    const double span = 20.0;
    for (int i = 0; i < n_samples; ++i) {
        const double t = span * (double(i) / (n_samples - 1.0)) - span / 2.0;
        out[i] = complex_t(1.0 / std::cosh(t / width) / width, 0.0);
    }
    return out;
}

// pulse::lorentzian (modulation.cpp:33-45): 1 / (1 + (2 x / fwhm)^2), x on a
// grid of step grid_step centred at sample n/2
ComplexVector pulse::lorentzian(const int& n_samples, const double& fwhm, const double& grid_step) {
    ComplexVector out(n_samples);
    for (int i = 0; i < n_samples; ++i) {
        const double x = grid_step * double(i - n_samples / 2);
        const double r = 2.0 * x / fwhm;
        out[i] = complex_t(1.0 / (1.0 + r * r), 0.0);
    }
    return out;
}

// pulse::rrc (modulation.cpp:47-74): the root-raised-cosine impulse
// response at t = i / osf - n/2 symbol periods, with the two removable
// singularities (t = 0 and |t| = 1 / (4 beta)) at their limits, multiplied
// by n / sum h^2 (n the number of samples), as the reference scales it
ComplexVector pulse::rrc(const int& n_symbols, const int& oversampling, const double& roll_off) {
    const int n = n_symbols * oversampling;
    const double b = roll_off;
    std::vector<double> h(n);
    double energy = 0;
    for (int i = 0; i < n; ++i) {
        const double t = double(i) / oversampling - n_symbols / 2.0;
        double v;
        if (t == 0) {
            v = 1.0 - b + 4.0 * b / math_pi;
        } else if (std::abs(t) == 1.0 / b / 4.0) {
            const double a = math_pi / (4.0 * b);
            v = b / std::sqrt(2.0) * ((1.0 + 2.0 / math_pi) * std::sin(a) + (1.0 - 2.0 / math_pi) * std::cos(a));
        } else {
            const double num = std::sin(math_pi * t * (1.0 - b)) + 4.0 * b * t * std::cos(math_pi * t * (1.0 + b));
            v = num / (1.0 - 16.0 * b * b * t * t) / (math_pi * t);
        }
        h[i] = v;
        energy += v * v;
    }
    ComplexVector out(n);
    const double scale = n / energy;
    for (int i = 0; i < n; ++i) out[i] = complex_t(h[i] * scale, 0.0);
    return out;
}

// random_16qam_symbols (modulation.cpp:76-83): every constellation point the
// same number of times (code i mod 16 at position i), then a uniform random
// permutation.  The reference shuffles with std::default_random_engine; this
// build draws the permutation from the runtime's Philox key (a Fisher-Yates
// shuffle over 64-bit draws), so realisations are reproducible by seed.
ComplexVector random_16qam_symbols(const int& n_symbols) {
    ComplexVector out(n_symbols);
    for (int i = 0; i < n_symbols; ++i) out[i] = gray_16qam((unsigned)(i % 16));
    // SplitMix64 stream keyed by the runtime seed and a per-call counter
    uint64_t state = noise_seed() ^ (0x9E3779B97F4A7C15ull * (1 + next_noise_stream()));
    auto draw = [&state]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int i = n_symbols - 1; i > 0; --i) {
        const int j = (int)(draw() % (uint64_t)(i + 1));
        std::swap(out[i], out[j]);
    }
    return out;
}

// rrc_shaping (modulation.cpp:86-90): upsample, circular RRC filter, x osf sqrt(gain)
Field rrc_shaping(const Field& symbols, const int& osf, const double& gain) {
    Domain d = symbols.domain();
    d.samples *= osf;
    d.time_step /= osf;
    Staged in(symbols);
    Staged out((size_t)d.size());
    check(focss_rrc_shaping(runtime(), FOCSS_F64, symbols.modes(), symbols.samples(), in.buf.ptr, osf, rrc_roll_off,
                            gain, out.buf.ptr),
          "focss_rrc_shaping");
    return out.fetch(d);
}

// rrc_sampling (modulation.cpp:92-96): circular RRC filter, every osf-th sample, x sqrt(gain)
Field rrc_sampling(const Field& signal, const int& osf, const double& gain) {
    assert(signal.samples() % osf == 0);
    Domain d = signal.domain();
    d.samples /= osf;
    d.time_step *= osf;
    Staged in(signal);
    Staged out((size_t)d.size());
    check(focss_rrc_sampling(runtime(), FOCSS_F64, signal.modes(), signal.samples(), in.buf.ptr, osf, rrc_roll_off,
                             gain, out.buf.ptr),
          "focss_rrc_sampling");
    return out.fetch(d);
}

// demodulate_16qam (modulation.cpp:98-110): the nearest constellation point
// by squared distance; of several at the same distance the lowest code
unsigned int demodulate_16qam(const complex_t& symbol) {
    unsigned best = 0;
    double best_d = std::norm(symbol - gray_16qam(0));
    for (unsigned code = 1; code < 16; ++code) {
        const double d = std::norm(symbol - gray_16qam(code));
        if (d < best_d) {
            best_d = d;
            best = code;
        }
    }
    return best;
}

// hard_decision_16qam (modulation.cpp:112-120): decide at unit power per
// mode, return the decided points at the signal's power
Field hard_decision_16qam(const Field& signal) {
    const double amp = std::sqrt(power_per_mode(signal));
    Field out(signal.domain());
    for (int i = 0; i < signal.size(); ++i) out[i] = amp * gray_16qam(demodulate_16qam(signal[i] / amp));
    return out;
}

// soft_decision_16qam (modulation.cpp:122-148): the hard decision, except
// that an axis within sqrt(10) * uncertainty of a decision boundary (0 or
// +-2/sqrt(10) at unit power) keeps the received coordinate
Field soft_decision_16qam(const Field& signal, const double& uncertainty) {
    const double amp = std::sqrt(power_per_mode(signal));
    const double boundary = 2.0 / std::sqrt(10.0);
    const double margin = std::sqrt(10.0) * uncertainty;
    auto near_boundary = [&](double v) {
        return std::abs(v) < margin || std::abs(v - boundary) < margin || std::abs(v + boundary) < margin;
    };
    Field out(signal.domain());
    for (int i = 0; i < signal.size(); ++i) {
        const complex_t u = signal[i] / amp;
        complex_t g = gray_16qam(demodulate_16qam(u));
        const double re = near_boundary(u.real()) ? u.real() : g.real();
        const double im = near_boundary(u.imag()) ? u.imag() : g.imag();
        out[i] = amp * complex_t(re, im);
    }
    return out;
}

// hd_sequence_16qam (modulation.cpp:150-158): the codes at unit power per mode
Vector<int> hd_sequence_16qam(const Field& signal) {
    const double to_unit = std::sqrt(signal.modes() / signal.average_power());
    Vector<int> codes(signal.size());
    for (int i = 0; i < signal.size(); ++i) codes[i] = (int)demodulate_16qam(to_unit * signal[i]);
    return codes;
}

namespace {
// errors between the decisions of tx and rx, both scaled by tx's unit-power
// factor, over samples [cut, samples - cut) of every mode; bits or symbols
unsigned count_errors(const Field& tx, const Field& rx, int cut, bool bits) {
    const double to_unit = std::sqrt(tx.modes() / tx.average_power());
    unsigned errors = 0;
    for (int m = 0; m < tx.modes(); ++m)
        for (int s = cut; s < tx.samples() - cut; ++s) {
            const unsigned a = demodulate_16qam(to_unit * tx(m, s));
            const unsigned b = demodulate_16qam(to_unit * rx(m, s));
            errors += bits ? bit_distance(a, b) : (a != b ? 1u : 0u);
        }
    return errors;
}
}  // namespace

// bit_error_rate_16qam (modulation.cpp:171-185): errors over 4 bits per symbol
double bit_error_rate_16qam(const Field& tx, const Field& rx) {
    assert(tx.size() == rx.size());
    return double(count_errors(tx, rx, 0, true)) / double(4 * tx.size());
}

// with an edge cut (modulation.cpp:187-202); the denominator subtracts 8 cut
// once, whatever the number of modes, as the reference does
double bit_error_rate_16qam(const Field& tx, const Field& rx, const int& cut) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    return double(count_errors(tx, rx, cut, true)) / double(4 * tx.size() - 8 * cut);
}

// symbol_error_rate_16qam (modulation.cpp:204-223); denominator size - modes cut
double symbol_error_rate_16qam(const Field& tx, const Field& rx, const int& cut) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    return double(count_errors(tx, rx, cut, false)) / double(tx.size() - tx.modes() * cut);
}

// wdm_mux (modulation.cpp:225-235): channel i moved to (i - n/2) x bandwidth, summed
Field wdm_mux(const Vector<Field>& channels, const double& channel_bandwidth) {
    assert(channels.size() > 0);
    const int n = channels.size();
    Field sum(channels.raw()[0].domain());
    for (int i = 0; i < n; ++i) sum += frequency_shift(channels.raw()[i], (i - n / 2) * channel_bandwidth);
    return sum;
}

// wdm_demux (modulation.cpp:237-246): every channel selected at baseband
Vector<Field> wdm_demux(const Field& field, const int& n_channels, const double& channel_bandwidth) {
    Vector<Field> out(n_channels, Field(field.domain()));
    for (int ch = 0; ch < n_channels; ++ch) out[ch] = wdm_select(field, ch - n_channels / 2, channel_bandwidth);
    return out;
}

// wdm_select (modulation.cpp:248-253): shift to baseband, then the
// brick-wall filter of half the channel bandwidth
Field wdm_select(const Field& field, const int& channel_index, const double& channel_bandwidth) {
    Staged in(field);
    Staged out((size_t)field.size());
    check(focss_wdm_select(runtime(), FOCSS_F64, field.modes(), field.samples(), field.dt(), channel_index,
                           channel_bandwidth, in.buf.ptr, out.buf.ptr),
          "focss_wdm_select");
    return out.fetch(field.domain());
}

// frequency_shift (modulation.cpp:255-264): x exp(i 2 pi f dt s)
Field frequency_shift(const Field& field, const double& frequency) {
    Staged dev(field);
    check(focss_frequency_shift(runtime(), FOCSS_F64, field.modes(), field.samples(), field.dt(), frequency,
                                dev.buf.ptr),
          "focss_frequency_shift");
    return dev.fetch(field.domain());
}

// bandpass_filter (modulation.cpp:266-277): zero the DFT bins
// cut .. N - cut, cut = floor(frequency x duration) + 1
Field bandpass_filter(const Field& field, const double& frequency) {
    Staged dev(field);
    check(focss_bandpass_filter(runtime(), FOCSS_F64, field.modes(), field.samples(), field.dt(), frequency,
                                dev.buf.ptr),
          "focss_bandpass_filter");
    return dev.fetch(field.domain());
}

}  // namespace focss
