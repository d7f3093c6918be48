#include "tasks/systems.h"

#include <cmath>
#include <iostream>

#include "focss/functions.h"
#include "focss/module/amplifier.h"
#include "focss/runtime.h"
#include "focss/solver/ssfm.h"
#include "tasks/parameters.h"

namespace focss {

// systems.cpp:10-15
LinkParameters reference_link() {
    LinkParameters p;
    p.fiber = {db_to_natural(::attenuation), disp_to_beta2(::dispersion, ::wavelength), ::nonlinearity,
               ::fiber_length, ::wavelength, ::polarization_coupling};
    p.spans = ::spans;
    p.noise_figure_db = ::noise_figure;
    p.max_phase_shift = ::max_phase_shift;
    p.max_step_size = ::max_forward_step;
    return p;
}

namespace {
double span_gain(const Fiber& f) { return std::exp(f.alpha * f.length); }  // db_to_linear(att * L)

int run_link(Field& signal, focss_link& l) {
    int steps = 0;
    run_link_program(signal, l, &steps);
    return steps;
}
}  // namespace

// per span: ADAPTIVE SSFM FORWARD, then amplify with ASE   systems.cpp:26-43
int forward_propagation(Field& signal, const LinkParameters& p) {
    SSFM ssfm(p.fiber);
    ssfm.set_grid(SSFM::ADAPTIVE);
    ssfm.set_maximum_phase_shift(p.max_phase_shift);
    ssfm.set_maximum_step_size(p.max_step_size);
    Amplifier amp(span_gain(p.fiber), db_to_linear(p.noise_figure_db));
    focss_link l = ssfm.link(signal, SSFM::FORWARD, p.spans);
    l.span_post_scale = std::sqrt(amp.gain());
    l.noise_sigma = amp.noise_factor() != 0 ? amp.noise_sigma(signal) : 0.0;
    // one Philox key per call; the span index selects the stream
    l.noise_seed = noise_seed() ^ (0x9E3779B97F4A7C15ull * (uint64_t)(next_noise_stream() + 1));
    return run_link(signal, l);
}

// per span: drop_power, then BACKWARD SSFM, UNIFORM(steps) or LOGARITHMIC
// when steps == 0                                   systems.cpp:73-91
void backward_propagation(Field& signal, const LinkParameters& p, const int& steps) {
    SSFM ssfm(p.fiber);
    if (steps == 0) {
        ssfm.set_grid(SSFM::LOGARITHMIC);
        ssfm.set_maximum_phase_shift(p.max_phase_shift);
        ssfm.set_maximum_step_size(p.max_step_size);
    } else {
        ssfm.set_grid(SSFM::UNIFORM);
        ssfm.set_total_steps(steps);
    }
    focss_link l = ssfm.link(signal, SSFM::BACKWARD, p.spans);
    double g = span_gain(p.fiber);
    l.span_pre_scale = g < 1 ? std::sqrt(g) : std::sqrt(1 / g);  // Amplifier::drop_power
    run_link(signal, l);
}

// SSFM(fiber.dispersive() * spans, 1) BACKWARD        systems.cpp:63-71
void cd_compensation(Field& signal, const LinkParameters& p, const int& n_spans) {
    SSFM ssfm(p.fiber.dispersive() * n_spans, 1);
    ssfm.run(signal, SSFM::BACKWARD);
}
}  // namespace focss

void back_to_back(focss::Field& signal) {
    focss::LinkParameters p = focss::reference_link();
    focss::Amplifier amp(focss::db_to_linear(::attenuation * ::fiber_length), focss::db_to_linear(::noise_figure));
    for (int i = 0; i < ::spans; ++i) amp.add_noise(signal);
    (void)p;
}

void forward_propagation(focss::Field& signal) {
    std::cout << ">> forward propagation" << std::endl;
    focss::forward_propagation(signal, focss::reference_link());
}

void forward_propagation(focss::Field& signal, const int& n_spans) {
    focss::LinkParameters p = focss::reference_link();
    p.spans = n_spans;
    focss::forward_propagation(signal, p);
}

void cd_compensation(focss::Field& signal, const int& n_spans) {
    focss::cd_compensation(signal, focss::reference_link(), n_spans);
}

void cd_compensation(focss::Field& signal) { focss::cd_compensation(signal, focss::reference_link(), ::spans); }

void backward_propagation(focss::Field& signal, const int& steps) {
    focss::backward_propagation(signal, focss::reference_link(), steps);
}
