#include "focss/equalizer/linear.h"

#include <cassert>

#include "focss/runtime.h"

namespace focss {
namespace {
// the two fields on the device, [modes][samples] complex128
struct FieldPair {
    DeviceBuffer d, a;
    FieldPair(const Field& desired, const Field& actual)
        : d(desired.raw(), (size_t)desired.size() * sizeof(complex_t)),
          a(actual.raw(), (size_t)actual.size() * sizeof(complex_t)) {
        assert(desired.domain() == actual.domain());
    }
};
}  // namespace

// phase_shift_equalizer.cpp:4-46
PhaseShiftEqualizer::PhaseShiftEqualizer() : angle_steps_(3600), estimated_angle_(0) {}
PhaseShiftEqualizer::PhaseShiftEqualizer(const int& angle_steps) : angle_steps_(angle_steps), estimated_angle_(0) {}
void PhaseShiftEqualizer::set_angle_steps(const int& angle_steps) { angle_steps_ = angle_steps; }
void PhaseShiftEqualizer::train(const Field& desired, const Field& actual) {
    estimated_angle_ = estimate(desired, actual);
}
double PhaseShiftEqualizer::get_angle() const { return estimated_angle_; }
Field PhaseShiftEqualizer::equalize(const Field& original) const { return original * i_exp(estimated_angle_); }
Field PhaseShiftEqualizer::train_equalize(const Field& desired, const Field& actual) const {
    return actual * i_exp(estimate(desired, actual));
}
double PhaseShiftEqualizer::estimate(const Field& desired, const Field& actual) const {
    FieldPair f(desired, actual);
    double angle = 0;
    check(focss_pse_estimate(runtime(), FOCSS_F64, actual.modes(), actual.samples(), f.d.ptr, f.a.ptr, angle_steps_,
                             &angle),
          "focss_pse_estimate");
    return angle;
}

// homothety_equalizer.cpp:4-33
HomothetyEqualizer::HomothetyEqualizer() : scalar_(1) {}
void HomothetyEqualizer::train(const Field& desired, const Field& actual) { scalar_ = estimate(desired, actual); }
Field HomothetyEqualizer::equalize(const Field& original) const { return scalar_ * original; }
Field HomothetyEqualizer::train_equalize(const Field& desired, const Field& actual) const {
    return estimate(desired, actual) * actual;
}
double HomothetyEqualizer::get_scalar() const { return scalar_; }
double HomothetyEqualizer::estimate(const Field& desired, const Field& actual) const {
    FieldPair f(desired, actual);
    double s = 0;
    check(focss_homothety_estimate(runtime(), FOCSS_F64, actual.modes(), actual.samples(), f.d.ptr, f.a.ptr, &s),
          "focss_homothety_estimate");
    return s;
}

// scalar_equalizer.cpp:4-33
ScalarEqualizer::ScalarEqualizer() : scalar_(1) {}
void ScalarEqualizer::train(const Field& desired, const Field& actual) { scalar_ = estimate(desired, actual); }
Field ScalarEqualizer::equalize(const Field& original) const { return scalar_ * original; }
Field ScalarEqualizer::train_equalize(const Field& desired, const Field& actual) const {
    return estimate(desired, actual) * actual;
}
complex_t ScalarEqualizer::get_scalar() const { return scalar_; }
complex_t ScalarEqualizer::estimate(const Field& desired, const Field& actual) const {
    FieldPair f(desired, actual);
    double s[2] = {0, 0};
    check(focss_scalar_estimate(runtime(), FOCSS_F64, actual.modes(), actual.samples(), f.d.ptr, f.a.ptr, s),
          "focss_scalar_estimate");
    return complex_t(s[0], s[1]);
}
}  // namespace focss
