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

// A copy of a field whose current data is on the device is made on the
// device; otherwise from the host copy.
Field::Field(const Field& o)
    : modes_(o.modes_), samples_(o.samples_), time_step_(o.time_step_), center_frequency_(o.center_frequency_),
      data_(o.data_.size()) {
    if (o.where_ == kDevice) {
        const size_t bytes = data_.size() * sizeof(complex_t);
        check(focss_device_alloc(runtime(), bytes, &dev_), "focss_device_alloc");
        check(focss_device_copy(runtime(), dev_, o.dev_, bytes), "focss_device_copy");
        where_ = kDevice;
    } else {
        data_ = o.data_;
        where_ = kHost;
    }
}

Field::Field(Field&& o) noexcept
    : modes_(o.modes_), samples_(o.samples_), time_step_(o.time_step_), center_frequency_(o.center_frequency_),
      data_(std::move(o.data_)), dev_(o.dev_), where_(o.where_) {
    o.dev_ = nullptr;
    o.where_ = kHost;
    o.data_.clear();
}

Field& Field::operator=(const Field& o) {
    if (this != &o) {
        Field t(o);
        *this = std::move(t);
    }
    return *this;
}

Field& Field::operator=(Field&& o) noexcept {
    if (this != &o) {
        release();
        modes_ = o.modes_;
        samples_ = o.samples_;
        time_step_ = o.time_step_;
        center_frequency_ = o.center_frequency_;
        data_ = std::move(o.data_);
        dev_ = o.dev_;
        where_ = o.where_;
        o.dev_ = nullptr;
        o.where_ = kHost;
        o.data_.clear();
    }
    return *this;
}

Field::~Field() { release(); }

void Field::release() {
    if (dev_) focss_device_free(runtime(), dev_);
    dev_ = nullptr;
}

const void* Field::device_data() const {
    const size_t bytes = data_.size() * sizeof(complex_t);
    if (!dev_) check(focss_device_alloc(runtime(), bytes, &dev_), "focss_device_alloc");
    if (where_ == kHost) {
        check(focss_memcpy_h2d(runtime(), dev_, data_.data(), bytes), "focss_memcpy_h2d");
        count_transfer();
        where_ = kBoth;
    }
    return dev_;
}

void* Field::device_data() {
    static_cast<const Field*>(this)->device_data();
    where_ = kDevice;
    return dev_;
}

void Field::to_host() const {
    if (where_ != kDevice) return;
    check(focss_memcpy_d2h(runtime(), data_.data(), dev_, data_.size() * sizeof(complex_t)), "focss_memcpy_d2h");
    count_transfer();
    where_ = kBoth;
}

ComplexVector Field::operator()(const int& mode) { return ComplexVector::proxy(samples_, raw() + samples_ * mode); }

// power / peak_power / average_power                  field.cpp:164-187
real_t Field::power(const int& sample) const {
    const complex_t* d = host_data();
    real_t p = 0;
    for (int i = sample; i < size(); i += samples_) p += std::norm(d[i]);
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
    const complex_t* d = host_data();
    real_t sum = 0;
    for (int i = 0; i < size(); ++i) sum += std::norm(d[i]);
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
    const complex_t* in = raw();
    for (int i = 0; i < size(); ++i) {
        *o = in[i];
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
    complex_t* o = out.raw();
    for (int i = 0; i < out.size(); ++i) {
        o[i] = *src;
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
    check(focss_decimate(runtime(), FOCSS_F64, modes_, samples_, factor, device_data(), out.device_data()),
          "focss_decimate");
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

// forward with x1/N, inverse unscaled, on the device copy  field.cpp:343-352
Field& Field::fft_inplace() {
    check(focss_fft(runtime(), FOCSS_F64, -1, modes_, samples_, device_data()), "focss_fft");
    return *this;
}

Field& Field::ifft_inplace() {
    check(focss_fft(runtime(), FOCSS_F64, +1, modes_, samples_, device_data()), "focss_fft");
    return *this;
}

Field Field::fft() const { return Field(*this).fft_inplace(); }
Field Field::ifft() const { return Field(*this).ifft_inplace(); }

// circular FFT filter with the impulse response centred at sample 0
// (field.cpp:358-372; an odd-length response loses its last tap, as in the
// reference); on the GPU (focss_filter)
Field& Field::filter_inplace(const ComplexVector& h) {
    check(focss_filter(runtime(), FOCSS_F64, modes_, samples_, h.size(), reinterpret_cast<const double*>(h.raw()),
                       device_data()),
          "focss_filter");
    return *this;
}

Field Field::filter(const ComplexVector& h) const { return Field(*this).filter_inplace(h); }

// by a scalar on the device while the device copy is the current one
// (a span loop's amplifier gain), else on the host     field.cpp:381-392
Field& Field::operator*=(const real_t& v) { return (*this) *= complex_t(v, 0); }
Field& Field::operator*=(const complex_t& v) {
    if (where_ == kDevice) {
        check(focss_field_scale(runtime(), FOCSS_F64, modes_, samples_, v.real(), v.imag(), device_data()),
              "focss_field_scale");
        return *this;
    }
    complex_t* d = host_data_mut();
    for (int i = 0; i < size(); ++i) d[i] *= v;
    return *this;
}
Field& Field::operator*=(const RealVector& v) {
    assert(v.size() == samples_);
    complex_t* d = host_data_mut();
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) d[s + m * samples_] *= v[s];
    return *this;
}
Field& Field::operator*=(const ComplexVector& v) {
    assert(v.size() == samples_);
    complex_t* d = host_data_mut();
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) d[s + m * samples_] *= v[s];
    return *this;
}
Field& Field::operator/=(const real_t& v) {
    complex_t* d = host_data_mut();
    for (int i = 0; i < size(); ++i) d[i] /= v;
    return *this;
}
Field& Field::operator/=(const complex_t& v) {
    complex_t* d = host_data_mut();
    for (int i = 0; i < size(); ++i) d[i] /= v;
    return *this;
}
Field& Field::operator/=(const RealVector& v) {
    assert(v.size() == samples_);
    complex_t* d = host_data_mut();
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) d[s + m * samples_] /= v[s];
    return *this;
}
Field& Field::operator/=(const ComplexVector& v) {
    assert(v.size() == samples_);
    complex_t* d = host_data_mut();
    for (int m = 0; m < modes_; ++m)
        for (int s = 0; s < samples_; ++s) d[s + m * samples_] /= v[s];
    return *this;
}
Field& Field::operator+=(const Field& o) {
    const complex_t* b = o.host_data();
    complex_t* d = host_data_mut();
    for (int i = 0; i < size(); ++i) d[i] += b[i];
    return *this;
}
Field& Field::operator-=(const Field& o) {
    const complex_t* b = o.host_data();
    complex_t* d = host_data_mut();
    for (int i = 0; i < size(); ++i) d[i] -= b[i];
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
