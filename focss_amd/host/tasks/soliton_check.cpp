// Known-answer checks of the split-step solver through the C++ API: two
// exact solutions of the lossless NLSE
//     dA/dz = -i (beta2/2) d^2A/dt^2 + i gamma |A|^2 A
// in the sign convention of SSFM::run (src/focss/solver/ssfm.cpp:142-172),
// in normalised units (gamma = kappa = 1, |beta2| = 1, one unit of length):
//   dark  (beta2 = +1): A = tanh(t + a) tanh(a - t) e^{i z}, two dark
//         solitons of unit background a = T/4 apart in the periodic window T,
//         continuous across the window edge and exact up to ~4 e^{-4a};
//   bright (beta2 = -1): A = sech(t) e^{i z / 2}.
// Each is propagated over z = 1 on the UNIFORM grid and compared with the
// analytic field; the result is the relative L2 error.
#include <cmath>
#include <iostream>

#include "focss/field.h"
#include "focss/functions.h"
#include "focss/solver/ssfm.h"
#include "tasks/soliton_check.h"

namespace focss {

namespace {
double rel_l2(const Field& a, const Field& b) {
    double num = 0, den = 0;
    for (int i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}

// the analytic field at distance z on n samples of the window [-T/2, T/2)
Field soliton(bool dark, int n, double window, double z) {
    const double dt = window / n;
    Field f(Domain{1, n, dt, 0.0});
    const double a = window / 4;
    for (int s = 0; s < n; ++s) {
        const double t = (s - n / 2) * dt;
        const double envelope = dark ? std::tanh(t + a) * std::tanh(a - t) : 1.0 / std::cosh(t);
        f[s] = envelope * i_exp(dark ? z : 0.5 * z);
    }
    return f;
}
}  // namespace

double soliton_check(bool dark, const int& n_samples, const int& n_steps, const double& window) {
    const Fiber unit = {0.0, dark ? 1.0 : -1.0, 1.0, 1.0, 0.0, 1.0};
    Field field = soliton(dark, n_samples, window, 0.0);
    SSFM solver(unit, n_steps);
    solver.set_grid(SSFM::UNIFORM);
    solver.run(field, SSFM::FORWARD);
    return rel_l2(field, soliton(dark, n_samples, window, unit.length));
}

}  // namespace focss
