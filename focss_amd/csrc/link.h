// Device-side step control for the SSFM span loop.
//
// One "link program" = n_spans x [span-start scale, SSFM::run over one fibre
// with the UNIFORM / LOGARITHMIC / ADAPTIVE grid, span-end scale, optional
// ASE].  It covers SSFM::run (src/focss/solver/ssfm.cpp:39-172),
// forward_propagation / backward_propagation / cd_compensation
// (src/tasks/systems.cpp:26-91) and the wdm_run.cpp:26-47 twins.
//
// The field alternates between a spectral multiply (every linear operator
// between two Kerr steps, merged: D(s1) D(s2) = D(s1 + s2)) and a time-domain
// operator (Kerr step, or span transition: ASE + next span's scale + the
// power reductions that drive the next step size).  plan_spectral() runs at
// the start of each spectral multiply and decides both that multiply and the
// following time-domain operator from the state and the reductions of the
// previous time-domain operator -- the data-dependent control of
// adaptive_solve / estimate_logarithmic_steps stays on the device.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define FOCSS_HD __host__ __device__
#define FOCSS_EXP ::exp
#define FOCSS_LOG ::log
#define FOCSS_SQRT ::sqrt
#else
#include <cmath>
#define FOCSS_HD
#define FOCSS_EXP std::exp
#define FOCSS_LOG std::log
#define FOCSS_SQRT std::sqrt
#endif

namespace focss_hip {

enum { GRID_UNIFORM = 0, GRID_LOGARITHMIC = 1, GRID_ADAPTIVE = 2 };
enum { PH_SPAN_START = 0, PH_IN_SPAN = 1, PH_FINAL_DONE = 2, PH_DONE = 3 };
enum { TOP_NONE = 0, TOP_NL = 1, TOP_TRANSITION = 2, TOP_EXIT = 3, TOP_ENTRY = 4 };

struct WaveStatic {
    double alpha, beta2, gamma, kappa, length;
    double kgamma;       // kappa * gamma, as multiplied in ssfm.cpp:161
    double abs_kgamma;   // |kappa * gamma|                   ssfm.cpp:137
    double dir;          // +1 FORWARD, -1 BACKWARD
    double phi_max, h_max;
    double pre_scale;    // at every span start (DBP drop_power)
    double post_scale;   // at every span end (amplify sqrt(G))
    double noise_sigma;  // per-quadrature ASE sigma at every span end, 0 = none
    double dw;           // 2 pi / (dt N)                     field.cpp:220
    unsigned long long seed;
    long long offset;    // element offset of the waveform in its batch buffer
    int grid, steps;     // steps = UNIFORM total_steps
    int n_spans, modes, samples;
    int linear_only;     // UNIFORM with kappa*gamma == 0: one multiply per span
    // ADAPTIVE step schedule: 1 = record each step taken into sched (its
    // length, negated for a span's last step), 2 = take the steps from sched
    // instead of estimating them; sched_cap entries
    int sched_mode, sched_cap;
    double* sched;
    // 1: hold the field's energy to an FP64 running expectation (complex64
    // programs; see plan_spectral)
    int energy_comp;
};

struct WaveState {
    int phase, span, step, n_log;
    int top;             // time-domain op to run next
    int noise_stream;    // ASE stream (span index) of that op, -1 = none
    int total_steps;     // split steps done over all spans
    int fault;           // 1: the field's power was not finite; the program stopped
    double d;            // distance done in the current span (plain double sum, ssfm.cpp:107)
    double h;            // |step| of the last planned step
    double pending;      // signed linear distance not yet applied
    double nl_coef;      // kappa gamma step for a TOP_NL op
    double t_scale;      // scale applied by a transition / entry op
    double log_inv_gain; // LOGARITHMIC: (e^{-alpha L} - 1) / n_log
    double e_ref;        // energy_comp: expected energy after the next time op
};

struct SpectralOp {
    int active;
    double s;     // signed linear distance of this multiply
    double gain;  // scalar gain (span scales), excluding loss and 1/N
    double loss;  // e^{-alpha s / 2}, the amplitude loss of the multiply
};

// Power reductions produced by each time-domain op: max and sum over samples
// of P(t) = sum_modes |A|^2 (Field::power / peak_power / average_power,
// field.cpp:164-187).
struct Reduction {
    unsigned long long peak_bits;  // non-negative double as bits (atomicMax)
    double sum;
    // energy_comp: the sum in fixed point, kFixOne = the expected energy
    // (integer adds are exact: the total does not depend on the order the
    // tiles finish in); kFixBad or more = a tile sum out of range
    unsigned long long fix;
};

constexpr double kFixOne = 1125899906842624.0;  // 2^50
constexpr unsigned long long kFixBad = 1ull << 62;

#ifdef __HIPCC__
// energy_comp: one tile's power sum into rd->fix, relative to the expected
// energy ex > 0, rounded to the nearest unit.  A tile sum above twice the
// expected total (or not finite) sets kFixBad instead; the adds of at most
// 1024 tiles below that stay under 2^61 and never reach its bit.
__device__ inline void add_fixed_sum(Reduction* rd, double sm, double ex) {
    const double q = sm * (kFixOne / ex);
    if (q >= 0 && q < 2 * kFixOne)
        atomicAdd(&rd->fix, (unsigned long long)(q + 0.5));
    else
        atomicOr(&rd->fix, kFixBad);
}
#endif

FOCSS_HD inline double bits_to_double(unsigned long long b) {
    union {
        unsigned long long u;
        double d;
    } c;
    c.u = b;
    return c.d;
}
FOCSS_HD inline unsigned long long double_to_bits(double d) {
    union {
        unsigned long long u;
        double d;
    } c;
    c.d = d;
    return c.u;
}

// estimate_next_adaptive_step (ssfm.cpp:136-140).  std::min(h_max, h) keeps
// h_max unless h < h_max, so a NaN peak gives h_max as in the reference.
FOCSS_HD inline double adaptive_h(const WaveStatic& w, double peak) {
    double phase_shift = w.abs_kgamma * peak;
    double h = w.phi_max / phase_shift;
    return (h < w.h_max) ? h : w.h_max;
}

// estimate_logarithmic_steps (ssfm.cpp:120-134)
FOCSS_HD inline int logarithmic_steps(const WaveStatic& w, double sum) {
    double avg = sum / w.samples;
    double l_nl = 1.0 / w.gamma / avg;
    double cl_nl_kappa = w.phi_max * l_nl / w.kappa;
    double l_eff = (1.0 - FOCSS_EXP(-w.alpha * w.length)) / w.alpha;
    double l_nl_eff = (1.0 - FOCSS_EXP(-w.alpha * cl_nl_kappa)) / w.alpha;
    int estimated = (int)(l_eff / l_nl_eff);
    int minimum = (int)(w.length / w.h_max);
    return minimum > estimated ? minimum : estimated;
}

// Take one split step of size h (unsigned): the multiply covers the pending
// half step plus the first half of this one, the Kerr op uses the full step.
FOCSS_HD inline void take_step(const WaveStatic& w, WaveState& st, SpectralOp& op, double h,
                               double step_signed, bool is_final) {
    op.s += st.pending + 0.5 * step_signed;
    st.pending = 0.5 * step_signed;
    st.h = h;
    st.nl_coef = w.kgamma * step_signed;
    st.top = TOP_NL;
    st.noise_stream = -1;
    st.step += 1;
    st.total_steps += 1;
    st.phase = is_final ? PH_FINAL_DONE : PH_IN_SPAN;
}

// One ADAPTIVE step (ssfm.cpp:92-118): h from the peak power, or the span's
// remainder when h would reach its end; with a replayed schedule, the
// recorded step instead (its sign marks the span's last step)
FOCSS_HD inline void adaptive_step(const WaveStatic& w, WaveState& st, SpectralOp& op, double peak) {
    if (w.sched_mode == 2 && st.total_steps < w.sched_cap) {
        const double v = w.sched[st.total_steps];
        const bool last = v < 0;
        const double h = last ? -v : v;
        take_step(w, st, op, h, w.dir * h, last);
        if (!last) st.d += h;
        return;
    }
    double h = adaptive_h(w, peak);
    if (st.d + h < w.length) {
        take_step(w, st, op, h, w.dir * h, false);
        st.d += h;
    } else {
        double hf = w.length - st.d;
        take_step(w, st, op, hf, w.dir * hf, true);
    }
}

// After a plan: the step just taken into the recorded schedule (one writer
// per waveform calls this)
FOCSS_HD inline void record_step(const WaveStatic& w, const WaveState& st) {
    if (w.sched_mode != 1 || w.grid != GRID_ADAPTIVE || st.top != TOP_NL) return;
    const int i = st.total_steps - 1;
    if (i >= 0 && i < w.sched_cap) w.sched[i] = st.phase == PH_FINAL_DONE ? -st.h : st.h;
}

// Plan the next spectral multiply and the time-domain op after it.
// red = reductions of the previous time-domain op.  Pure function of its
// inputs, so every workgroup of a pass computes the same plan.
FOCSS_HD inline SpectralOp plan_step(const WaveStatic& w, WaveState& st, double peak, double sum) {
    SpectralOp op;
    op.active = 0;
    op.s = 0;
    op.gain = 1;
    if (st.phase == PH_DONE) {
        st.top = TOP_NONE;
        return op;
    }
    // a NaN or infinite sample makes the power sum non-finite (the reference
    // propagates it silently; std::min(h_max, NaN) even keeps stepping,
    // ssfm.cpp:139): stop the program and report it
    if (!(sum <= 1.7976931348623157e308)) {
        st.fault = 1;
        st.phase = PH_DONE;
        st.top = TOP_NONE;
        return op;
    }
    op.active = 1;
    for (int guard = 0; guard < 2 * w.n_spans + 4; ++guard) {
        if (st.phase == PH_SPAN_START) {
            st.d = 0;
            st.step = 0;
            if (w.linear_only) {
                // fibre.dispersive(): Kerr is the identity; the whole span is
                // one linear operator of length L (ssfm.cpp:62-72 with gamma = 0)
                op.s += st.pending + w.dir * w.length;
                st.pending = 0;
                st.phase = PH_FINAL_DONE;
                continue;
            }
            if (w.grid == GRID_UNIFORM) {
                double h = w.length / w.steps;  // ssfm.cpp:63-64
                take_step(w, st, op, h, w.dir * h, w.steps == 1);
            } else if (w.grid == GRID_LOGARITHMIC) {
                st.n_log = logarithmic_steps(w, sum);
                double inv_gain = FOCSS_EXP(-w.alpha * w.length) - 1;  // ssfm.cpp:77-78
                inv_gain /= st.n_log;
                st.log_inv_gain = inv_gain;
                double inv_alpha = -1.0 / w.alpha * w.dir;
                double step = inv_alpha * FOCSS_LOG(1 + inv_gain / (1 + inv_gain * 0));
                take_step(w, st, op, step * w.dir, step, st.n_log == 1);
            } else {
                adaptive_step(w, st, op, peak);
            }
            return op;
        }
        if (st.phase == PH_IN_SPAN) {
            if (w.grid == GRID_UNIFORM) {
                double h = w.length / w.steps;
                take_step(w, st, op, h, w.dir * h, st.step + 1 == w.steps);
            } else if (w.grid == GRID_LOGARITHMIC) {
                double inv_alpha = -1.0 / w.alpha * w.dir;
                double g = st.log_inv_gain;
                double step = inv_alpha * FOCSS_LOG(1 + g / (1 + g * st.step));
                take_step(w, st, op, step * w.dir, step, st.step + 1 == st.n_log);
            } else {
                // next_step was estimated after the previous Kerr op (ssfm.cpp:102-108)
                adaptive_step(w, st, op, peak);
            }
            return op;
        }
        // PH_FINAL_DONE: end of span
        op.s += st.pending;
        st.pending = 0;
        op.gain *= w.post_scale;
        int finished = st.span;
        st.span += 1;
        int noise = (w.noise_sigma != 0) ? finished : -1;
        if (st.span >= w.n_spans) {
            st.top = TOP_EXIT;
            st.noise_stream = noise;
            st.t_scale = 1;
            st.phase = PH_DONE;
            return op;
        }
        if (noise >= 0 || w.grid != GRID_UNIFORM) {
            st.top = TOP_TRANSITION;
            st.noise_stream = noise;
            st.t_scale = w.pre_scale;
            st.phase = PH_SPAN_START;
            return op;
        }
        // UNIFORM without ASE: next span's drop_power is a scalar, merge it
        op.gain *= w.pre_scale;
        st.phase = PH_SPAN_START;
    }
    return op;
}

// plan_step, and for energy_comp programs the compensated energy: the
// multiply, the inverse DFT and the Kerr op change the energy
// sum_t P(t) by exactly gain^2 e^{-alpha s} (dispersion and Kerr phases have
// modulus 1), and a span transition without ASE by t_scale^2, so the energy
// after those ops is known in FP64 from the one measured after the entry or
// the last transition with ASE (ASE makes that one random).  Complex64
// rounding loses a steady ~2e-7 of it per split step, which also lowers
// every later Kerr phase; the multiply after each such op takes
// sqrt(expected / measured) into its scalar, so the drift does not
// accumulate over the steps and spans.  sum must be run-to-run
// deterministic (formed in tile order) for the plans to be.
//
// expected_energy: what the time op that st plans leaves the field with
// when its arithmetic is exact; 0 when that is not known (entry, ASE).
FOCSS_HD inline double expected_energy(const WaveState& st) {
    if (st.top == TOP_NL) return st.e_ref;
    if (st.top == TOP_TRANSITION && st.noise_stream < 0) return st.e_ref * st.t_scale * st.t_scale;
    return 0;
}

FOCSS_HD inline SpectralOp plan_spectral(const WaveStatic& w, WaveState& st, double peak,
                                         double sum) {
    const double expect = w.energy_comp ? expected_energy(st) : 0;  // of the op just run
    SpectralOp op = plan_step(w, st, peak, sum);
    op.loss = op.active ? FOCSS_EXP(-0.5 * w.alpha * op.s) : 1.0;
    if (!w.energy_comp || !op.active) return op;
    const bool held = expect > 0 && sum > 0;
    const double base = held ? expect : sum;
    st.e_ref = base * (op.gain * op.loss) * (op.gain * op.loss);
    if (held) op.gain *= FOCSS_SQRT(base / sum);
    return op;
}

// energy_comp: the measured energy from the fixed-point sum when the
// expected one is known (a tile sum out of range: the float sum, whose
// order of adds varies), else from the tile sums in tile order (ordered(),
// which the time op writes then only: once a span at most)
template <typename F>
FOCSS_HD inline double comp_sum(const WaveState& st, const Reduction& rd, F ordered) {
    const double ex = expected_energy(st);
    if (!(ex > 0)) return ordered();
    return rd.fix < kFixBad ? (double)rd.fix * (ex / kFixOne) : rd.sum;
}

// Initial state: the first time-domain op applies span 0's scale and
// measures the field for the first step estimate.
FOCSS_HD inline WaveState initial_state(const WaveStatic& w) {
    WaveState st;
    st.phase = PH_SPAN_START;
    st.span = 0;
    st.step = 0;
    st.n_log = 0;
    st.top = TOP_ENTRY;
    st.noise_stream = -1;
    st.total_steps = 0;
    st.fault = 0;
    st.d = 0;
    st.h = 0;
    st.pending = 0;
    st.nl_coef = 0;
    st.t_scale = w.pre_scale;
    st.log_inv_gain = 0;
    st.e_ref = 0;
    return st;
}

}  // namespace focss_hip
