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

// AWGN realisation on the field's device copy: Philox stream = the
// runtime's next stream number under the runtime's seed  amplifier.cpp:20-28
void Amplifier::add_noise(Field& field) const {
    check(focss_add_noise(runtime(), FOCSS_F64, field.modes(), field.samples(), noise_sigma(field), noise_seed(),
                          next_noise_stream(), field.device_data()),
          "focss_add_noise");
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
