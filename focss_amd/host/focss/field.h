// focss::Domain / focss::Field -- the signal container of the focss C++ API
// with the reference's interface (src/focss/field.h:20-162).  Storage is the
// reference's layout: mode-major complex<double>, (mode, sample) at
// data[sample + mode * samples] (src/focss/field.cpp:140-155).  The data
// stays in host memory as in the reference; transforms run on the GPU
// through the focss C-ABI (include/focss.h).
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
    Field(const Field&) = default;
    Field(Field&&) = default;
    Field& operator=(const Field&) = default;
    Field& operator=(Field&&) = default;
    ~Field() = default;

    complex_t operator[](const int& index) const { return data_[index]; }
    complex_t operator()(const int& mode, const int& sample) const { return data_[sample + mode * samples_]; }
    complex_t& operator[](const int& index) { return data_[index]; }
    complex_t& operator()(const int& mode, const int& sample) { return data_[sample + mode * samples_]; }
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

    complex_t* raw() { return data_.data(); }
    const complex_t* raw() const { return data_.data(); }

  private:
    int modes_;
    int samples_;
    real_t time_step_;
    real_t center_frequency_;
    std::vector<complex_t> data_;
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
