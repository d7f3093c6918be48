// The compensated energy of the step planner (focss_amd/csrc/link.h),
// compiled for the host.  Each program is driven through plan_spectral on a
// scalar model of the field's energy: a multiply scales it by
// (gain loss)^2 and then loses `delta` (what complex64 rounding does on the
// device), a Kerr op keeps it, a span transition scales it by t_scale^2 and
// adds the ASE's energy.  Beside it runs the same program without the loss
// and without the compensation, whose energy is exact.  Prints one line per
// case: name, E / E_exact - 1, Kerr steps, largest gain difference of the
// uncompensated planner against plan_step.
#include <cmath>
#include <cstdio>

#include "link.h"

using namespace focss_hip;

namespace {

WaveStatic program(int grid, int steps, int spans, double dir, double pre, double post, double noise, int comp) {
    WaveStatic w{};
    w.alpha = 4.6e-5;  // 0.2 dB/km
    w.beta2 = -2.17e-26;
    w.gamma = 1.3e-3;
    w.kappa = 1.0;
    w.length = 80e3;
    w.kgamma = w.kappa * w.gamma;
    w.abs_kgamma = std::fabs(w.kgamma);
    w.dir = dir;
    w.phi_max = 1e-3;
    w.h_max = 1e3;
    w.pre_scale = pre;
    w.post_scale = post;
    w.noise_sigma = noise;
    w.dw = 1e9;
    w.grid = grid;
    w.steps = steps;
    w.n_spans = spans;
    w.modes = 1;
    w.samples = 1 << 12;
    w.energy_comp = comp;
    return w;
}

void run(const char* name, WaveStatic w, double delta) {
    WaveStatic wr = w;
    wr.energy_comp = 0;
    WaveState st = initial_state(w), ref = initial_state(wr), bare = initial_state(w);
    double e = 2.0 * w.pre_scale * w.pre_scale, er = e;  // after the entry op
    double gain_diff = 0;
    unsigned noise_seq = 12345;
    int kerr = 0;
    for (int it = 0; it < 1000000; ++it) {
        const double peak = 2e-2 * e / 2.0;  // the ADAPTIVE grid's step follows the power
        const SpectralOp op = plan_spectral(w, st, peak, e);
        const SpectralOp opr = plan_spectral(wr, ref, 2e-2 * er / 2.0, er);
        if (!w.energy_comp) {
            const SpectralOp opb = plan_step(w, bare, peak, e);
            gain_diff = std::fmax(gain_diff, std::fabs(op.gain - opb.gain));
        }
        if (op.active != opr.active) {
            std::printf("%s mismatch 0 0 0\n", name);
            return;
        }
        if (!op.active) break;
        e *= (op.gain * op.loss) * (op.gain * op.loss) * (1 - delta);
        er *= (opr.gain * opr.loss) * (opr.gain * opr.loss);
        if (st.top == TOP_NL) ++kerr;
        if (st.top == TOP_TRANSITION) {
            noise_seq = noise_seq * 1103515245u + 12345u;
            const double ase = st.noise_stream >= 0 ? 0.1 * (noise_seq % 1000) / 1000.0 : 0.0;
            e = e * st.t_scale * st.t_scale + ase;
            er = er * ref.t_scale * ref.t_scale + ase;
        }
    }
    std::printf("%s %.6e %d %.3e\n", name, e / er - 1, kerr, gain_diff);
}

}  // namespace

int main() {
    const double g = std::sqrt(std::exp(4.6e-5 * 80e3));  // amplitude gain undoing one span's loss
    const double d = 3e-7;
    run("uniform_comp", program(GRID_UNIFORM, 100, 8, 1, 1, g, 0, 1), d);
    run("uniform_plain", program(GRID_UNIFORM, 100, 8, 1, 1, g, 0, 0), d);
    run("adaptive_comp", program(GRID_ADAPTIVE, 0, 8, 1, 1, g, 0, 1), d);
    run("adaptive_plain", program(GRID_ADAPTIVE, 0, 8, 1, 1, g, 0, 0), d);
    run("logarithmic_comp", program(GRID_LOGARITHMIC, 0, 8, 1, 1, g, 0, 1), d);
    run("uniform_ase_comp", program(GRID_UNIFORM, 100, 8, 1, 1, g, 1e-3, 1), d);
    run("dbp_comp", program(GRID_UNIFORM, 50, 8, -1, 1 / g, 1, 0, 1), d);
    run("dbp_plain", program(GRID_UNIFORM, 50, 8, -1, 1 / g, 1, 0, 0), d);
    // the fixed-point sum: what the tiles add, read back by comp_sum
    WaveState st{};
    st.top = TOP_NL;
    st.e_ref = 3.25;
    Reduction rd{};
    const double measured = 3.2499876543;
    rd.fix = (unsigned long long)(measured * (kFixOne / st.e_ref) + 0.5);
    rd.sum = 1.0;
    const double fixed = comp_sum(st, rd, [] { return -1.0; });
    rd.fix = kFixBad | 17;
    const double bad = comp_sum(st, rd, [] { return -1.0; });
    st.top = TOP_ENTRY;
    const double reset = comp_sum(st, rd, [] { return -1.0; });
    std::printf("fixed %.6e %.1f %.1f\n", fixed / measured - 1, bad, reset);
    return 0;
}
