// focss::SSFM -- split-step Fourier solver with the reference's interface
// (src/focss/solver/ssfm.h:8-54).  run() hands the field to the device as a
// one-span link program (focss_link_run_host): every FFT, linear step, Kerr
// step and the adaptive / logarithmic step control execute on the GPU.
#ifndef FOCSS_SOLVER_SSFM_H_
#define FOCSS_SOLVER_SSFM_H_

#include "focss.h"
#include "focss/field.h"
#include "focss/module/fiber.h"

namespace focss {
class SSFM {
  public:
    typedef const int Grid;
    static const int UNIFORM;
    static const int LOGARITHMIC;
    static const int ADAPTIVE;

    typedef const double Direction;
    static const double FORWARD;
    static const double BACKWARD;

    SSFM();
    SSFM(const Fiber& fiber);
    SSFM(const Fiber& fiber, const int& total_steps);

    void set_fiber(const Fiber& fiber);
    void set_grid(const Grid& grid);
    void set_total_steps(const int& total_steps);
    void set_maximum_phase_shift(const double& maximum_phase_shift);
    void set_maximum_step_size(const double& maximum_step_size);

    void run(Field& field, const Direction& direction = FORWARD);

    // the device link program run() would execute for `field`
    focss_link link(const Field& field, const double& direction, const int& n_spans) const;
    int last_steps() const { return last_steps_; }

  private:
    Fiber fiber_;
    int grid_;
    int total_steps_;
    double maximum_phase_shift_;
    double maximum_step_size_;
    int last_steps_;
};

// Runs one link program on `field` at the runtime's precision: FP64 on the
// field's device copy (which stays current), FP32 through the host copy.
void run_link_program(Field& field, const focss_link& link, int* total_steps);
}  // namespace focss

#endif  // FOCSS_SOLVER_SSFM_H_
