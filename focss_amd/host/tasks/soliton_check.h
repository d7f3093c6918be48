// Known-answer checks of the split-step solver (tasks/soliton_check.cpp).
#ifndef FOCSS_TASKS_SOLITON_CHECK_H_
#define FOCSS_TASKS_SOLITON_CHECK_H_

namespace focss {
// Relative L2 error after z = 1 of the dark-soliton pair (dark = true,
// beta2 = +1) or the bright sech soliton (beta2 = -1), on n_samples points
// of a periodic window of `window` normalised time units, with n_steps
// UNIFORM split steps.
double soliton_check(bool dark, const int& n_samples, const int& n_steps, const double& window = 40.0);
}  // namespace focss

#endif  // FOCSS_TASKS_SOLITON_CHECK_H_
