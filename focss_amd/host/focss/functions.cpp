#include "focss/functions.h"

#include <cassert>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

#include "focss/runtime.h"

namespace focss {

void setup() { reseed_global_urng(); }

void reseed_global_urng() {
    reseed_global_urng((uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count());
}

void reseed_global_urng(uint64_t seed) { set_noise_seed(seed); }

// evm2 / q2                                          functions.cpp:27-80
double evm2_factor(const ComplexVector& tx, const ComplexVector& rx) {
    assert(tx.size() == rx.size());
    double num = 0, den = 0;
    for (int i = 0; i < tx.size(); ++i) {
        num += std::norm(tx[i] - rx[i]);
        den += std::norm(tx[i]);
    }
    return num / den;
}

double evm2_factor(const Field& tx, const Field& rx) {
    assert(tx.size() == rx.size());
    double num = 0, den = 0;
    for (int i = 0; i < tx.size(); ++i) {
        num += std::norm(tx[i] - rx[i]);
        den += std::norm(tx[i]);
    }
    return num / den;
}

double evm2_factor(const Field& tx, const Field& rx, const int& cut) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    double num = 0, den = 0;
    for (int m = 0; m < tx.modes(); ++m)
        for (int s = cut; s < tx.samples() - cut; ++s) {
            num += std::norm(tx(m, s) - rx(m, s));
            den += std::norm(tx(m, s));
        }
    return num / den;
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

// CSV transmission files: per sample tx modes then rx modes as re,im
// (functions.cpp:174-221).  Written with 17 significant digits so doubles
// round-trip (the reference's precision(15) does not).
void save_transmission(std::ostream& out, const Field& tx, const Field& rx) {
    assert(tx.modes() == rx.modes() && tx.samples() == rx.samples());
    for (int i = 0; i < tx.samples(); ++i) {
        for (int m = 0; m < tx.modes(); ++m) out << tx(m, i).real() << ',' << tx(m, i).imag() << ',';
        for (int m = 0; m < rx.modes() - 1; ++m) out << rx(m, i).real() << ',' << rx(m, i).imag() << ',';
        out << rx(rx.modes() - 1, i).real() << ',' << rx(rx.modes() - 1, i).imag() << '\n';
    }
}

void load_transmission(std::istream& in, Field* tx, Field* rx) {
    assert(tx->modes() == rx->modes() && tx->samples() == rx->samples());
    double re, im;
    char sep = ',';
    for (int i = 0; i < tx->samples(); ++i) {
        in >> re >> sep >> im;
        (*tx)(0, i) = complex_t(re, im);
        for (int m = 1; m < tx->modes(); ++m) {
            in >> sep >> re >> sep >> im;
            (*tx)(m, i) = complex_t(re, im);
        }
        for (int m = 0; m < tx->modes(); ++m) {
            in >> sep >> re >> sep >> im;
            (*rx)(m, i) = complex_t(re, im);
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
