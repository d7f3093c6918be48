#include "focss/solver/ssfm.h"

#include <iostream>

#include "focss/runtime.h"

namespace focss {
const int SSFM::UNIFORM = 0;
const int SSFM::LOGARITHMIC = 1;
const int SSFM::ADAPTIVE = 2;
const double SSFM::FORWARD = 1.0;
const double SSFM::BACKWARD = -1.0;

// the reference leaves grid and limits uninitialised in SSFM() (ssfm.cpp:16);
// here they start as UNIFORM with no steps, which the device rejects
SSFM::SSFM() : fiber_{0, 0, 0, 0, 0, 1}, grid_(UNIFORM), total_steps_(0), maximum_phase_shift_(0),
               maximum_step_size_(0), last_steps_(0) {}
SSFM::SSFM(const Fiber& f) : SSFM() { fiber_ = f; }
SSFM::SSFM(const Fiber& f, const int& steps) : SSFM() {
    fiber_ = f;
    grid_ = UNIFORM;
    total_steps_ = steps;
}

void SSFM::set_fiber(const Fiber& f) { fiber_ = f; }
void SSFM::set_grid(const Grid& g) { grid_ = g; }
void SSFM::set_total_steps(const int& n) { total_steps_ = n; }
void SSFM::set_maximum_phase_shift(const double& p) { maximum_phase_shift_ = p; }
void SSFM::set_maximum_step_size(const double& h) { maximum_step_size_ = h; }

focss_link SSFM::link(const Field& field, const double& direction, const int& n_spans) const {
    focss_link l = {};
    l.fiber = {fiber_.alpha, fiber_.beta2, fiber_.gamma, fiber_.length, fiber_.wavelength, fiber_.kappa};
    l.grid = {grid_, total_steps_, maximum_phase_shift_, maximum_step_size_};
    l.direction = direction;
    l.n_spans = n_spans;
    l.modes = field.modes();
    l.samples = field.samples();
    l.dt = field.dt();
    l.center_frequency = field.center_frequency();
    l.span_pre_scale = 1;
    l.span_post_scale = 1;
    l.noise_sigma = 0;
    l.noise_seed = 0;
    return l;
}

// ssfm.cpp:39-60 (grids 62-118) on the device
void SSFM::run(Field& field, const Direction& direction) {
    focss_link l = link(field, direction, 1);
    run_link_program(field, l, &last_steps_);
    if (grid_ == ADAPTIVE) std::cout << "    * SSFM done " << last_steps_ << " steps" << std::endl;
}

// FP64 programs run on the field's device copy, which stays current; FP32
// programs convert through the host copy (focss_link_run_host)
void run_link_program(Field& field, const focss_link& l, int* steps) {
    if (precision() == FOCSS_F64) {
        const int64_t offset = 0;
        check(focss_link_run(runtime(), FOCSS_F64, 1, &l, field.device_data(), &offset, steps), "focss_link_run");
    } else {
        check(focss_link_run_host(runtime(), precision(), &l, reinterpret_cast<double*>(field.raw()), steps),
              "focss_link_run_host");
    }
}
}  // namespace focss
