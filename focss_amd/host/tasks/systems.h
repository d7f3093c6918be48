// Link-level drivers with the reference's signatures (src/tasks/systems.h:6-11)
// plus overloads taking explicit link parameters.  Each driver hands the
// whole span loop to the device as one link program (focss_link_run_host):
// gain, drop_power and ASE are fused into the SSFM passes.
#ifndef TASKS_SYSTEMS_H_
#define TASKS_SYSTEMS_H_

#include "focss/field.h"
#include "focss/module/fiber.h"

void back_to_back(focss::Field& signal);
void cd_compensation(focss::Field& signal);
void cd_compensation(focss::Field& signal, const int& spans);
void forward_propagation(focss::Field& signal);
void forward_propagation(focss::Field& signal, const int& spans);
void backward_propagation(focss::Field& signal, const int& steps = 0);

namespace focss {
struct LinkParameters {
    Fiber fiber;              // one span
    int spans;
    double noise_figure_db;   // EDFA noise figure, forward links
    double max_phase_shift;   // ADAPTIVE / LOGARITHMIC
    double max_step_size;     // [m]
};
LinkParameters reference_link();  // the constants of tasks/parameters.h

int forward_propagation(Field& signal, const LinkParameters& p);  // returns split steps
void backward_propagation(Field& signal, const LinkParameters& p, const int& steps);
void cd_compensation(Field& signal, const LinkParameters& p, const int& spans);
}  // namespace focss

#endif  // TASKS_SYSTEMS_H_
