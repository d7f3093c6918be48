#include "focss/functions.h"

#include <cassert>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "focss/runtime.h"

namespace focss {

void setup() { reseed_global_urng(); }

void reseed_global_urng() {
    reseed_global_urng((uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count());
}

void reseed_global_urng(uint64_t seed) { set_noise_seed(seed); }

// evm2 / q2 (functions.cpp:27-80): the error energy sum |tx - rx|^2 over
// the reference energy sum |tx|^2, on every element or, with an edge cut,
// on samples [cut, samples - cut) of every mode
namespace {
struct Energies {
    double error = 0, reference = 0;
    void add(const complex_t& t, const complex_t& r) {
        error += std::norm(t - r);
        reference += std::norm(t);
    }
    double ratio() const { return error / reference; }
};
}  // namespace

double evm2_factor(const ComplexVector& tx, const ComplexVector& rx) {
    assert(tx.size() == rx.size());
    Energies e;
    const complex_t* t = tx.raw();
    const complex_t* r = rx.raw();
    for (int i = 0; i < tx.size(); ++i) e.add(t[i], r[i]);
    return e.ratio();
}

double evm2_factor(const Field& tx, const Field& rx) {
    assert(tx.size() == rx.size());
    Energies e;
    for (int i = 0; i < tx.size(); ++i) e.add(tx.raw()[i], rx.raw()[i]);
    return e.ratio();
}

double evm2_factor(const Field& tx, const Field& rx, const int& cut) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    Energies e;
    const int n = tx.samples();
    for (int m = 0; m < tx.modes(); ++m) {
        const complex_t* t = tx.raw() + (size_t)m * n;
        const complex_t* r = rx.raw() + (size_t)m * n;
        for (int s = cut; s < n - cut; ++s) e.add(t[s], r[s]);
    }
    return e.ratio();
}

double q2_factor(const ComplexVector& tx, const ComplexVector& rx) { return -10 * std::log10(evm2_factor(tx, rx)); }
double q2_factor(const Field& tx, const Field& rx) { return -10 * std::log10(evm2_factor(tx, rx)); }
double q2_factor(const Field& tx, const Field& rx, const int& cut) {
    return -10 * std::log10(evm2_factor(tx, rx, cut));
}

// conversions                                         functions.cpp:82-96
double db_to_linear(const double& db) { return std::pow(10, db / 10); }
double dbm_to_watts(const double& dbm) { return 1e-3 * std::pow(10, dbm / 10); }
double db_to_natural(const double& db) { return db * std::log(10) / 10; }
double disp_to_beta2(const double& d, const double& wl) { return -wl * wl * d / (2 * math_pi * light_speed); }

double sinc(const double& x) {
    if (x == 0) return 1;
    return std::sin(math_pi * x) / math_pi / x;
}

complex_t i_exp(const double& x) { return complex_t(std::cos(x), std::sin(x)); }

// single-sequence transforms (functions.cpp:108-169), on the device
void fft_inplace(const int& size, complex_t* io) { device_fft(-1, 1, size, io); }
void ifft_inplace(const int& size, complex_t* io) { device_fft(+1, 1, size, io); }
void fft(const int& size, complex_t* in, complex_t* out) {
    if (out != in) std::memcpy(out, in, sizeof(complex_t) * size);
    device_fft(-1, 1, size, out);
}
void ifft(const int& size, complex_t* in, complex_t* out) {
    if (out != in) std::memcpy(out, in, sizeof(complex_t) * size);
    device_fft(+1, 1, size, out);
}
void fft_inplace(ComplexVector* d) { fft_inplace(d->size(), d->raw()); }
void ifft_inplace(ComplexVector* d) { ifft_inplace(d->size(), d->raw()); }
ComplexVector fft(const ComplexVector& d) {
    ComplexVector c(d);
    fft_inplace(&c);
    return c;
}
ComplexVector ifft(const ComplexVector& d) {
    ComplexVector c(d);
    ifft_inplace(&c);
    return c;
}

// CSV transmission files (functions.cpp:174-221): one line per sample,
// the real and imaginary parts of tx mode 0 .. M-1, then of rx mode
// 0 .. M-1, separated by commas.  Written with 17 significant digits so
// doubles round-trip (the reference's precision(15) does not).
void save_transmission(std::ostream& out, const Field& tx, const Field& rx) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    const int modes = tx.modes();
    for (int s = 0; s < tx.samples(); ++s) {
        for (int k = 0; k < 2 * modes; ++k) {
            const complex_t v = k < modes ? tx(k, s) : rx(k - modes, s);
            out << v.real() << ',' << v.imag() << (k + 1 < 2 * modes ? ',' : '\n');
        }
    }
}

// reads tx->samples() lines of 4 M comma-separated numbers
void load_transmission(std::istream& in, Field* tx, Field* rx) {
    assert(tx->modes() == rx->modes() && tx->samples() == rx->samples());
    const int modes = tx->modes();
    std::vector<double> row(4 * (size_t)modes);
    std::string line;
    for (int s = 0; s < tx->samples(); ++s) {
        if (!std::getline(in, line)) throw std::runtime_error("load_transmission: file ends early");
        const char* p = line.c_str();
        for (double& v : row) {
            char* end = nullptr;
            v = std::strtod(p, &end);
            if (end == p) throw std::runtime_error("load_transmission: malformed line");
            p = (*end == ',') ? end + 1 : end;
        }
        for (int m = 0; m < modes; ++m) {
            (*tx)(m, s) = complex_t(row[2 * m], row[2 * m + 1]);
            (*rx)(m, s) = complex_t(row[2 * (modes + m)], row[2 * (modes + m) + 1]);
        }
    }
}

void save_transmission(const std::string& filename, const Field& tx, const Field& rx) {
    std::ofstream file(filename);
    file.precision(std::numeric_limits<double>::max_digits10);
    save_transmission(file, tx, rx);
}

void load_transmission(const std::string& filename, Field* tx, Field* rx) {
    std::ifstream file(filename);
    load_transmission(file, tx, rx);
}
}  // namespace focss
