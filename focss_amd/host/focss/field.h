// focss::Domain / focss::Field -- the signal container of the focss C++ API
// with the reference's interface (src/focss/field.h:20-162).  Storage is the
// reference's layout: mode-major complex<double>, (mode, sample) at
// data[sample + mode * samples] (src/focss/field.cpp:140-155).
//
// A Field holds a host copy and, once a device operation has touched it, a
// device copy, each current or stale.  The operations that run on the GPU
// (SSFM::run, fft_inplace / ifft_inplace, filter_inplace, decimate,
// Amplifier::add_noise, and *= by a scalar while the device copy is the
// current one) work on the device copy and leave the host copy stale; any
// host access (operator[], operator(), raw(), the power and arithmetic
// members) first brings the host copy up to date, and non-const host access
// marks the device copy stale.  A span loop of device operations therefore
// crosses PCIe once each way, however many spans it runs.  A pointer from
// raw() stays valid for host writes until the next device operation.
#ifndef FOCSS_FIELD_H_
#define FOCSS_FIELD_H_

#include <vector>

#include "focss/definitions.h"

namespace focss {

struct Domain {
    int modes;
    int samples;
    real_t time_step;
    real_t center_frequency;

    inline int size() const { return modes * samples; }
    inline bool operator==(const Domain& o) const {
        return modes == o.modes && samples == o.samples && time_step == o.time_step &&
               center_frequency == o.center_frequency;
    }
};

class Field {
  public:
    Field();
    explicit Field(const Domain& domain);
    explicit Field(const Domain& domain, const ComplexVector& data);
    Field(const Field&);
    Field(Field&&) noexcept;
    Field& operator=(const Field&);
    Field& operator=(Field&&) noexcept;
    ~Field();

    complex_t operator[](const int& index) const { return host_data()[index]; }
    complex_t operator()(const int& mode, const int& sample) const { return host_data()[sample + mode * samples_]; }
    complex_t& operator[](const int& index) { return host_data_mut()[index]; }
    complex_t& operator()(const int& mode, const int& sample) { return host_data_mut()[sample + mode * samples_]; }
    ComplexVector operator()(const int& mode);

    real_t power(const int& sample) const;
    real_t peak_power() const;
    real_t average_power() const;
    Field& peak_normalize(const real_t& power = 1);
    Field& average_normalize(const real_t& power = 1);

    int size() const { return modes_ * samples_; }
    int modes() const { return modes_; }
    int samples() const { return samples_; }
    real_t duration() const { return samples_ * time_step_; }
    real_t bandwidth() const { return 1.0 / time_step_; }
    real_t time_step() const { return time_step_; }
    real_t sampling_rate() const { return 1.0 / time_step_; }
    real_t center_frequency() const { return center_frequency_; }
    real_t dt() const { return time_step_; }
    real_t df() const { return 1.0 / (time_step_ * samples_); }
    real_t dw() const { return 2.0 * math_pi / (time_step_ * samples_); }
    real_t t(const int& sample) const { return static_cast<real_t>(sample) * dt(); }
    real_t f(const int& sample) const { return center_frequency_ + (sample - samples_ / 2) * df(); }
    real_t w(const int& sample) const { return 2 * math_pi * f(sample); }
    Domain domain() const { return {modes_, samples_, time_step_, center_frequency_}; }

    RealVector time_grid() const;
    RealVector frequency_grid() const;
    RealVector temporal_power() const;
    RealVector spectral_power() const;

    Field upsample(const int& factor) const;
    Field downsample(const int& factor) const;
    Field decimate(const int& factor) const;

    Field& fft_shift();
    Field& fft_inplace();
    Field& ifft_inplace();
    Field fft() const;
    Field ifft() const;
    Field& filter_inplace(const ComplexVector& impulse_response);
    Field filter(const ComplexVector& impulse_response) const;

    Field& operator*=(const real_t&);
    Field& operator*=(const complex_t&);
    Field& operator*=(const RealVector&);
    Field& operator*=(const ComplexVector&);
    Field& operator/=(const real_t&);
    Field& operator/=(const complex_t&);
    Field& operator/=(const RealVector&);
    Field& operator/=(const ComplexVector&);
    Field& operator+=(const Field&);
    Field& operator-=(const Field&);

    complex_t* raw() { return host_data_mut(); }
    const complex_t* raw() const { return host_data(); }

    // Device residency (no counterpart in the reference's interface): the
    // device copy (complex<double>, the same layout), current on return.
    // The non-const form is for writing on the device: it marks the host
    // copy stale.  to_host() brings the host copy up to date.
    void* device_data();
    const void* device_data() const;
    void to_host() const;
    bool host_current() const { return where_ != kDevice; }
    bool device_current() const { return where_ != kHost; }

  private:
    enum { kHost = 0, kDevice = 1, kBoth = 2 };  // which copies are current
    const complex_t* host_data() const {
        if (where_ == kDevice) to_host();
        return data_.data();
    }
    complex_t* host_data_mut() {
        if (where_ == kDevice) to_host();
        where_ = kHost;
        return data_.data();
    }
    void release();

    int modes_;
    int samples_;
    real_t time_step_;
    real_t center_frequency_;
    mutable std::vector<complex_t> data_;
    mutable void* dev_ = nullptr;
    mutable int where_ = kHost;
};

Field operator*(const Field&, const real_t&);
Field operator*(const Field&, const complex_t&);
Field operator*(const Field&, const RealVector&);
Field operator*(const Field&, const ComplexVector&);
Field operator*(const real_t&, const Field&);
Field operator*(const complex_t&, const Field&);
Field operator*(const RealVector&, const Field&);
Field operator*(const ComplexVector&, const Field&);
Field operator/(const Field&, const real_t&);
Field operator/(const Field&, const complex_t&);
Field operator/(const Field&, const RealVector&);
Field operator/(const Field&, const ComplexVector&);
Field operator+(const Field&, const Field&);
Field operator-(const Field&, const Field&);

// batched device DFT of `modes` rows of `samples` (reference normalisation)
void device_fft(int sign, int modes, int samples, complex_t* data);
}  // namespace focss

#endif  // FOCSS_FIELD_H_
