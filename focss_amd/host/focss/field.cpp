#include "focss/field.h"

#include <algorithm>
#include <cassert>
#include <cmath>

#include "focss/functions.h"
#include "focss/runtime.h"

namespace focss {

Field::Field() : modes_(0), samples_(0), time_step_(0), center_frequency_(0) {}

Field::Field(const Domain& d)
    : modes_(d.modes), samples_(d.samples), time_step_(d.time_step), center_frequency_(d.center_frequency),
      data_((size_t)d.modes * d.samples, complex_t(0, 0)) {
    assert(d.modes > 0 && d.samples > 0 && d.time_step > 0);
}

Field::Field(const Domain& d, const ComplexVector& v)
    : modes_(d.modes), samples_(d.samples), time_step_(d.time_step), center_frequency_(d.center_frequency),
      data_(v.raw(), v.raw() + v.size()) {
    assert(d.size() == v.size() && d.time_step > 0);
}

ComplexVector Field::operator()(const int& mode) { return ComplexVector::proxy(samples_, raw() + samples_ * mode); }

// power / peak_power / average_power                  field.cpp:164-187
real_t Field::power(const int& sample) const {
    real_t p = 0;
    for (int i = sample; i < size(); i += samples_) p += std::norm(data_[i]);
    return p;
}

real_t Field::peak_power() const {
    real_t peak = 0;
    for (int s = 0; s < samples_; ++s) {
        real_t p = power(s);
        if (peak < p) peak = p;
    }
    return peak;
}

real_t Field::average_power() const {
    real_t sum = 0;
    for (const complex_t& v : data_) sum += std::norm(v);
    return sum / samples_;
}

Field& Field::peak_normalize(const real_t& p) { return (*this) *= std::sqrt(p / peak_power()); }
Field& Field::average_normalize(const real_t& p) { return (*this) *= std::sqrt(p / average_power()); }

RealVector Field::time_grid() const {
    RealVector g(samples_);
    for (int s = 0; s < samples_; ++s) g[s] = t(s);
    return g;
}

RealVector Field::frequency_grid() const {
    RealVector g(samples_);
    for (int s = 0; s < samples_; ++s) g[s] = f(s);
    return g;
}

RealVector Field::temporal_power() const {
    RealVector g(samples_);
    for (int s = 0; s < samples_; ++s) g[s] = power(s);
    return g;
}

RealVector Field::spectral_power() const {
    RealVector g(samples_);
    Field spectrum = fft().fft_shift();
    for (int s = 0; s < samples_; ++s) g[s] = spectrum.power(s);
    return g;
}

// upsample / downsample act on the flat mode-major array, as the reference
// does (field.cpp:272-297): upsampling puts value i at flat position
// factor i (zeros between), downsampling keeps every factor-th value of the
// flat array
Field Field::upsample(const int& factor) const {
    Domain d = domain();
    d.samples *= factor;
    d.time_step /= factor;
    Field out(d);
    complex_t* o = out.raw();
    for (const complex_t& v : data_) {
        *o = v;
        o += factor;
    }
    return out;
}

Field Field::downsample(const int& factor) const {
    assert(samples_ % factor == 0);
    Domain d = domain();
    d.samples /= factor;
    d.time_step *= factor;
    Field out(d);
    const complex_t* src = raw();
    for (complex_t& v : out.data_) {
        v = *src;
        src += factor;
    }
    return out;
}

// decimate (field.cpp:299-320): the lowest N/factor bins of the spectrum
// (the first N/(2 factor) non-negative ones and the rest from the top), back
// to the time domain at the lower rate; on the GPU (focss_decimate)
Field Field::decimate(const int& factor) const {
    assert(samples_ % factor == 0);
    Domain d = domain();
    d.samples /= factor;
    d.time_step *= factor;
    Field out(d);
    DeviceBuffer in(raw(), sizeof(complex_t) * (size_t)size());
    DeviceBuffer res(sizeof(complex_t) * (size_t)out.size());
    check(focss_decimate(runtime(), FOCSS_F64, modes_, samples_, factor, in.ptr, res.ptr), "focss_decimate");
    res.download(out.raw());
    return out;
}

// swap the two halves of every mode (field.cpp:322-330)
Field& Field::fft_shift() {
    const int half = samples_ / 2;
    for (int m = 0; m < modes_; ++m) {
        complex_t* row = raw() + (size_t)m * samples_;
        std::swap_ranges(row, row + half, row + half);
    }
    return *this;
}

void device_fft(int sign, int modes, int samples, complex_t* data) {
    focss_ctx* ctx = runtime();
    size_t bytes = (size_t)modes * samples * sizeof(complex_t);
    void* d = nullptr;
    check(focss_device_alloc(ctx, bytes, &d), "focss_device_alloc");
    check(focss_memcpy_h2d(ctx, d, data, bytes), "focss_memcpy_h2d");
    check(focss_fft(ctx, FOCSS_F64, sign, modes, samples, d), "focss_fft");
    check(focss_memcpy_d2h(ctx, data, d, bytes), "focss_memcpy_d2h");
    check(focss_device_free(ctx, d), "focss_device_free");
}

// forward with x1/N, inverse unscaled                   field.cpp:343-352
Field& Field::fft_inplace() {
    device_fft(-1, modes_, samples_, raw());
    return *this;
}

Field& Field::ifft_inplace() {
    device_fft(+1, modes_, samples_, raw());
    return *this;
}

Field Field::fft() const { return Field(*this).fft_inplace(); }
Field Field::ifft() const { return Field(*this).ifft_inplace(); }

// circular FFT filter with the impulse response centred at sample 0
// (field.cpp:358-372; an odd-length response loses its last tap, as in the
// reference); on the GPU (focss_filter)
Field& Field::filter_inplace(const ComplexVector& h) {
    DeviceBuffer dev(raw(), sizeof(complex_t) * (size_t)size());
    check(focss_filter(runtime(), FOCSS_F64, modes_, samples_, h.size(), reinterpret_cast<const double*>(h.raw()),
                       dev.ptr),
          "focss_filter");
    dev.download(raw());
    return *this;
}

Field Field::filter(const ComplexVector& h) const { return Field(*this).filter_inplace(h); }

Field& Field::operator*=(const real_t& v) {
    for (auto& x : data_) x *= v;
    return *this;
}
Field& Field::operator*=(const complex_t& v) {
    for (auto& x : data_) x *= v;
    return *this;
}
Field& Field::operator*=(const RealVector& v) {
    assert(v.size() == samples_);
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) data_[s + m * samples_] *= v[s];
    return *this;
}
Field& Field::operator*=(const ComplexVector& v) {
    assert(v.size() == samples_);
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) data_[s + m * samples_] *= v[s];
    return *this;
}
Field& Field::operator/=(const real_t& v) {
    for (auto& x : data_) x /= v;
    return *this;
}
Field& Field::operator/=(const complex_t& v) {
    for (auto& x : data_) x /= v;
    return *this;
}
Field& Field::operator/=(const RealVector& v) {
    assert(v.size() == samples_);
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) data_[s + m * samples_] /= v[s];
    return *this;
}
Field& Field::operator/=(const ComplexVector& v) {
    assert(v.size() == samples_);
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) data_[s + m * samples_] /= v[s];
    return *this;
}
Field& Field::operator+=(const Field& o) {
    for (int i = 0; i < size(); ++i) data_[i] += o.data_[i];
    return *this;
}
Field& Field::operator-=(const Field& o) {
    for (int i = 0; i < size(); ++i) data_[i] -= o.data_[i];
    return *this;
}

Field operator*(const Field& f, const real_t& v) { return Field(f) *= v; }
Field operator*(const Field& f, const complex_t& v) { return Field(f) *= v; }
Field operator*(const Field& f, const RealVector& v) { return Field(f) *= v; }
Field operator*(const Field& f, const ComplexVector& v) { return Field(f) *= v; }
Field operator*(const real_t& v, const Field& f) { return Field(f) *= v; }
Field operator*(const complex_t& v, const Field& f) { return Field(f) *= v; }
Field operator*(const RealVector& v, const Field& f) { return Field(f) *= v; }
Field operator*(const ComplexVector& v, const Field& f) { return Field(f) *= v; }
Field operator/(const Field& f, const real_t& v) { return Field(f) /= v; }
Field operator/(const Field& f, const complex_t& v) { return Field(f) /= v; }
Field operator/(const Field& f, const RealVector& v) { return Field(f) /= v; }
Field operator/(const Field& f, const ComplexVector& v) { return Field(f) /= v; }
Field operator+(const Field& a, const Field& b) { return Field(a) += b; }
Field operator-(const Field& a, const Field& b) { return Field(a) -= b; }
}  // namespace focss
