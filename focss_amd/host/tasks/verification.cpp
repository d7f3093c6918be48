// Dark-soliton check of the SSFM (src/tasks/verification.cpp:6-52): UNIFORM
// FORWARD over fibre {alpha 0, beta2 1, gamma 1, L 1, kappa 1} against
// sqrt(P) e^{i P z} tanh(sqrt(P) t) on a 10-unit window; prints the relative
// L2 error.
#include <cmath>
#include <iostream>

#include "focss/field.h"
#include "focss/functions.h"
#include "focss/solver/ssfm.h"

using namespace focss;

namespace {
const Fiber verification_fiber = {0.0, 1.0, 1.0, 1.0, 0.0, 1.0};
const double time_domain = 10.0;
const double soliton_power = 1.0;

Field dark_pulse(const int& points, const double& z) {
    double dt = time_domain / static_cast<double>(points);
    Field field(Domain{1, points, dt, 0.0});
    double a = std::sqrt(soliton_power);
    for (int i = 0; i < field.size(); ++i) {
        double t = static_cast<double>(i - field.size() / 2) * dt;
        field[i] = a * i_exp(soliton_power * z) * std::tanh(a * t);
    }
    return field;
}
}  // namespace

double verification(const int& n_samples, const int& n_steps) {
    Field field = dark_pulse(n_samples, 0);
    Field solution = dark_pulse(n_samples, verification_fiber.length);
    SSFM ssfm;
    ssfm.set_fiber(verification_fiber);
    ssfm.set_grid(SSFM::UNIFORM);
    ssfm.set_total_steps(n_steps);
    ssfm.run(field, SSFM::FORWARD);
    double error = std::sqrt((field - solution).average_power() / solution.average_power());
    std::cout << error << std::endl;
    return error;
}
