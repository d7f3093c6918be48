// focss::Amplifier -- EDFA gain and ASE (src/focss/module/amplifier.h:7-22).
#ifndef FOCSS_MODULE_AMPLIFIER_H_
#define FOCSS_MODULE_AMPLIFIER_H_

#include "focss/field.h"

namespace focss {
class Amplifier {
  public:
    Amplifier();
    Amplifier(const double& gain);
    Amplifier(const double& gain, const double& noise_factor);

    void amplify(Field& field) const;
    void add_noise(Field& field) const;
    void give_power(Field& field) const;
    void drop_power(Field& field) const;

    double gain() const { return gain_; }
    double noise_factor() const { return noise_factor_; }
    // per-quadrature ASE standard deviation for a field (amplifier.cpp:21-25)
    double noise_sigma(const Field& field) const;

  private:
    double gain_;
    double noise_factor_;
};
}  // namespace focss

#endif  // FOCSS_MODULE_AMPLIFIER_H_
