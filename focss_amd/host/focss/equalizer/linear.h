// focss::PhaseShiftEqualizer, HomothetyEqualizer, ScalarEqualizer -- the
// constant equalisers with the reference's interface
// (src/focss/equalizer/linear.h:17-64).  Each estimate is one device pass
// over the two fields (focss_pse_estimate / focss_homothety_estimate /
// focss_scalar_estimate); applying the constant is a host Field product, as
// in the reference.
#ifndef FOCSS_EQUALIZER_LINEAR_H_
#define FOCSS_EQUALIZER_LINEAR_H_

#include "focss/field.h"
#include "focss/functions.h"

namespace focss {

class PhaseShiftEqualizer {
  public:
    PhaseShiftEqualizer();
    explicit PhaseShiftEqualizer(const int& angle_steps);

    void train(const Field& desired, const Field& actual);
    Field equalize(const Field& original) const;
    Field train_equalize(const Field& desired, const Field& actual) const;

    double get_angle() const;
    void set_angle_steps(const int& angle_steps);

  private:
    int angle_steps_;
    double estimated_angle_;

    double estimate(const Field& desired, const Field& actual) const;
};

class HomothetyEqualizer {
  public:
    HomothetyEqualizer();

    void train(const Field& desired, const Field& actual);
    Field equalize(const Field& original) const;
    Field train_equalize(const Field& desired, const Field& actual) const;

    double get_scalar() const;

  private:
    double scalar_;
    double estimate(const Field& desired, const Field& actual) const;
};

class ScalarEqualizer {
  public:
    ScalarEqualizer();

    void train(const Field& desired, const Field& actual);
    Field equalize(const Field& original) const;
    Field train_equalize(const Field& desired, const Field& actual) const;

    complex_t get_scalar() const;

  private:
    complex_t scalar_;
    complex_t estimate(const Field& desired, const Field& actual) const;
};
}  // namespace focss

#endif  // FOCSS_EQUALIZER_LINEAR_H_
