#include "focss/module/amplifier.h"

#include <cmath>

#include "focss/runtime.h"

namespace focss {
Amplifier::Amplifier() : gain_(1), noise_factor_(0) {}
Amplifier::Amplifier(const double& gain) : gain_(gain), noise_factor_(0) {}
Amplifier::Amplifier(const double& gain, const double& nf) : gain_(gain), noise_factor_(nf) {}

// x sqrt(G), then ASE when NF != 0                 amplifier.cpp:15-18
void Amplifier::amplify(Field& field) const {
    field *= std::sqrt(gain_);
    if (noise_factor_ != 0) add_noise(field);
}

double Amplifier::noise_sigma(const Field& field) const {
    double variance = planck * field.center_frequency();
    variance *= noise_factor_ * (gain_ - 1) / 2;
    variance *= field.bandwidth();
    return std::sqrt(variance / 2);
}

// AWGN realisation on the device: Philox stream = the runtime's next
// stream number under the runtime's seed            amplifier.cpp:20-28
void Amplifier::add_noise(Field& field) const {
    focss_ctx* ctx = runtime();
    size_t bytes = (size_t)field.size() * sizeof(complex_t);
    void* d = nullptr;
    check(focss_device_alloc(ctx, bytes, &d), "focss_device_alloc");
    check(focss_memcpy_h2d(ctx, d, field.raw(), bytes), "focss_memcpy_h2d");
    check(focss_add_noise(ctx, FOCSS_F64, field.modes(), field.samples(), noise_sigma(field), noise_seed(),
                          next_noise_stream(), d),
          "focss_add_noise");
    check(focss_memcpy_d2h(ctx, field.raw(), d, bytes), "focss_memcpy_d2h");
    check(focss_device_free(ctx, d), "focss_device_free");
}

void Amplifier::give_power(Field& field) const {
    if (gain_ > 1)
        field *= std::sqrt(gain_);
    else
        field *= std::sqrt(1 / gain_);
}

void Amplifier::drop_power(Field& field) const {
    if (gain_ < 1)
        field *= std::sqrt(gain_);
    else
        field *= std::sqrt(1 / gain_);
}
}  // namespace focss
