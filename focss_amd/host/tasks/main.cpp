// focss_tasks: command-line driver over the focss C++ API.
//   focss_tasks soliton dark|bright <samples> <steps>   (prints the relative L2 error)
//   focss_tasks dbp     <in> <out> <samples> <modes> <dt> <spans> <steps> <alpha> <beta2> <gamma> <length> <kappa>
//   focss_tasks cdc     <in> <out> <samples> <modes> <dt> <spans> <alpha> <beta2> <gamma> <length> <kappa>
//   focss_tasks sino    <tx> <rx> <out> <samples> <modes> <radius> <type> <lw_re> <lw_im>
//                       (out: weights (2R+1)^2, intercept, linear weight, then
//                        the equalized field (rx - cubic(tx)) / w, complex<double>)
//   focss_tasks wdmppe  <tx> <rx> <out> <samples> <c_radius> <d_radius>
//                       (tx, rx: 6 envelope rows; out: C (2c+1)^2 if c >= 0,
//                        D (2d+1)^2 if d >= 0, intercept, linear weight, then
//                        the equalized COI (rx_coi - cubic(tx)) / w, 2 rows)
//   focss_tasks linear-eq <desired> <actual> <out> <samples> <modes> <angle_steps>
//                       (out: PSE angle, homothety scalar, complex scalar as 4
//                        doubles, then PhaseShiftEqualizer::train_equalize)
//   focss_tasks forward <in> <out> <samples> <modes> <dt> <spans> <alpha> <beta2> <gamma> <length> <kappa>
//                       <max_phase_shift> <max_step> <nf_db> <seed> <center_frequency>
//   focss_tasks save-field <in> <set> <tag> <index> <channels> <modes> <samples>
//                       (in: the channels' fields one after another)
//   focss_tasks load-field <set> <tag> <index> <out>   (prints the header)
//   focss_tasks modulation <in> <out> <samples> <modes> <dt> <osf> <channel_bw>
//                       (out, one after another: decimate(osf), filter by
//                        pulse::rrc(samples/osf, osf, 0.25), rrc_sampling(osf),
//                        rrc_shaping(rrc_sampling), wdm_select(1, bw),
//                        frequency_shift(bw/3), bandpass_filter(bw/2),
//                        upsample(2), downsample(2), fft_shift,
//                        hard_decision_16qam(rrc_sampling), wdm_mux of
//                        {f, f/2, f/4} at bw)
//   focss_tasks save-transmission <tx> <rx> <csv> <modes> <samples>
//   focss_tasks load-transmission <csv> <tx> <rx> <modes> <samples>
// <in>/<out>/<tx>/<rx> are raw little-endian complex<double> files in the
// field layout.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "focss/equalizer/linear.h"
#include "focss/equalizer/sino.h"
#include "focss/equalizer/wdm_sino.h"
#include "focss/field.h"
#include "focss/functions.h"
#include "focss/module/amplifier.h"
#include "focss/processing/modulation.h"
#include "focss/runtime.h"
#include "tasks/field_io.h"
#include "focss/solver/ssfm.h"
#include "tasks/soliton_check.h"
#include "tasks/systems.h"

namespace {
focss::Field read_field(const char* path, int samples, int modes, double dt, double fc) {
    focss::Field f(focss::Domain{modes, samples, dt, fc});
    std::ifstream in(path, std::ios::binary);
    in.read(reinterpret_cast<char*>(f.raw()), (std::streamsize)sizeof(focss::complex_t) * f.size());
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    return f;
}

void write_field(const char* path, const focss::Field& f) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(f.raw()), (std::streamsize)sizeof(focss::complex_t) * f.size());
}

focss::LinkParameters params(char** a) {
    focss::LinkParameters p = focss::reference_link();
    p.fiber.alpha = std::atof(a[0]);
    p.fiber.beta2 = std::atof(a[1]);
    p.fiber.gamma = std::atof(a[2]);
    p.fiber.length = std::atof(a[3]);
    p.fiber.kappa = std::atof(a[4]);
    return p;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: focss_tasks soliton|dbp|cdc|forward ...\n");
        return 2;
    }
    std::string cmd = argv[1];
    if (cmd == "soliton" && argc == 5) {
        double e = focss::soliton_check(std::strcmp(argv[2], "dark") == 0, std::atoi(argv[3]), std::atoi(argv[4]));
        std::cout.precision(17);
        std::cout << e << std::endl;
        return e < 1 ? 0 : 1;
    }
    if ((cmd == "dbp" && argc == 14) || (cmd == "cdc" && argc == 13) || (cmd == "forward" && argc == 18)) {
        int samples = std::atoi(argv[4]), modes = std::atoi(argv[5]);
        double dt = std::atof(argv[6]);
        int spans = std::atoi(argv[7]);
        double fc = cmd == "forward" ? std::atof(argv[17]) : 0.0;
        focss::Field f = read_field(argv[2], samples, modes, dt, fc);
        if (cmd == "dbp") {
            focss::LinkParameters p = params(argv + 9);
            p.spans = spans;
            focss::backward_propagation(f, p, std::atoi(argv[8]));
        } else if (cmd == "cdc") {
            focss::LinkParameters p = params(argv + 8);
            focss::cd_compensation(f, p, spans);
        } else {
            focss::LinkParameters p = params(argv + 8);
            p.spans = spans;
            p.max_phase_shift = std::atof(argv[13]);
            p.max_step_size = std::atof(argv[14]);
            p.noise_figure_db = std::atof(argv[15]);
            focss::reseed_global_urng(std::strtoull(argv[16], nullptr, 0));
            int steps = focss::forward_propagation(f, p);
            std::cout << "split steps " << steps << std::endl;
        }
        write_field(argv[3], f);
        return 0;
    }
    // the reference's per-span calling pattern (src/tasks/wdm_run.cpp:26-47):
    // for each span SSFM::run, then Amplifier::amplify with ASE; prints the
    // host<->device copies of the field the loop made
    if (cmd == "span-loop" && argc == 17) {
        int samples = std::atoi(argv[4]), modes = std::atoi(argv[5]);
        double dt = std::atof(argv[6]);
        int spans = std::atoi(argv[7]), steps = std::atoi(argv[8]);
        focss::Field f = read_field(argv[2], samples, modes, dt, std::atof(argv[16]));
        focss::LinkParameters p = params(argv + 9);
        focss::reseed_global_urng(std::strtoull(argv[15], nullptr, 0));
        focss::SSFM ssfm(p.fiber, steps);
        focss::Amplifier amp(std::exp(p.fiber.alpha * p.fiber.length), focss::db_to_linear(std::atof(argv[14])));
        const uint64_t before = focss::transfer_count();
        for (int s = 0; s < spans; ++s) {
            ssfm.run(f, focss::SSFM::FORWARD);
            amp.amplify(f);
        }
        f.to_host();
        std::cout << "transfers " << (focss::transfer_count() - before) << std::endl;
        write_field(argv[3], f);
        return 0;
    }
    if (cmd == "sino" && argc == 11) {
        int samples = std::atoi(argv[5]), modes = std::atoi(argv[6]), radius = std::atoi(argv[7]);
        focss::Field tx = read_field(argv[2], samples, modes, 1e-11, 0.0);
        focss::Field rx = read_field(argv[3], samples, modes, 1e-11, 0.0);
        focss::SINO sino(radius, 1.0, (focss::SinoType)std::atoi(argv[8]));
        sino.set_linear_weight(focss::complex_t(std::atof(argv[9]), std::atof(argv[10])));
        sino.train(tx, rx);
        std::ofstream out(argv[4], std::ios::binary);
        out.write(reinterpret_cast<const char*>(sino.weights().data()),
                  (std::streamsize)(sino.weights().size() * sizeof(focss::complex_t)));
        focss::complex_t ic = sino.get_intercept(), lw = sino.get_linear_weight();
        out.write(reinterpret_cast<const char*>(&ic), sizeof(ic));
        out.write(reinterpret_cast<const char*>(&lw), sizeof(lw));
        focss::Field eq = sino.equalize(rx, tx);
        out.write(reinterpret_cast<const char*>(eq.raw()), (std::streamsize)(eq.size() * sizeof(focss::complex_t)));
        return 0;
    }
    if (cmd == "linear-eq" && argc == 8) {
        int samples = std::atoi(argv[5]), modes = std::atoi(argv[6]);
        focss::Field d = read_field(argv[2], samples, modes, 1e-11, 0.0);
        focss::Field a = read_field(argv[3], samples, modes, 1e-11, 0.0);
        focss::PhaseShiftEqualizer pse(std::atoi(argv[7]));
        pse.train(d, a);
        focss::HomothetyEqualizer hom;
        hom.train(d, a);
        focss::ScalarEqualizer sca;
        sca.train(d, a);
        const double head[4] = {pse.get_angle(), hom.get_scalar(), sca.get_scalar().real(),
                                sca.get_scalar().imag()};
        std::ofstream out(argv[4], std::ios::binary);
        out.write(reinterpret_cast<const char*>(head), sizeof(head));
        focss::Field eq = pse.train_equalize(d, a);
        out.write(reinterpret_cast<const char*>(eq.raw()), (std::streamsize)(eq.size() * sizeof(focss::complex_t)));
        return 0;
    }
    if (cmd == "wdmppe" && argc == 8) {
        int samples = std::atoi(argv[5]);
        focss::Field tx = read_field(argv[2], samples, 6, 1e-11, 0.0);
        focss::Field rx = read_field(argv[3], samples, 6, 1e-11, 0.0);
        focss::WdmPPE ppe(std::atoi(argv[6]), std::atoi(argv[7]));
        ppe.train(tx, rx);
        std::ofstream out(argv[4], std::ios::binary);
        auto put = [&](const focss::complex_t* p, size_t n) {
            out.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(focss::complex_t)));
        };
        put(ppe.c_weights().data(), ppe.c_weights().size());
        put(ppe.d_weights().data(), ppe.d_weights().size());
        focss::complex_t ic = ppe.get_intercept(), lw = ppe.get_linear_weight();
        put(&ic, 1);
        put(&lw, 1);
        focss::Field eq = ppe.equalize(tx, rx);
        put(eq.raw(), (size_t)eq.size());
        return 0;
    }
    if (cmd == "save-field" && argc == 9) {
        int channels = std::atoi(argv[6]), modes = std::atoi(argv[7]), samples = std::atoi(argv[8]);
        focss::Field all = read_field(argv[2], samples, channels * modes, 1.0, 0.0);
        focss::Vector<focss::Field> data(channels);
        for (int ch = 0; ch < channels; ++ch) {
            focss::Field f(focss::Domain{modes, samples, 1.0, 0.0});
            for (int m = 0; m < modes; ++m)
                for (int s = 0; s < samples; ++s) f(m, s) = all(ch * modes + m, s);
            data[ch] = f;
        }
        save_field(data, argv[3], argv[4], std::atoi(argv[5]));
        return 0;
    }
    if (cmd == "load-field" && argc == 6) {
        focss::FieldFileHeader h{};
        std::ifstream in(focss::field_path(argv[2], argv[3], std::atoi(argv[4])), std::ios::binary);
        if (!in) {
            std::fprintf(stderr, "cannot open field file\n");
            return 2;
        }
        focss::Field f;
        try {
            f = focss::load_field(in, &h);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
        std::cout << h.channels << ' ' << h.samples << ' ' << h.modes << ' ' << h.flags << std::endl;
        write_field(argv[5], f);
        return 0;
    }
    if (cmd == "modulation" && argc == 9) {
        int samples = std::atoi(argv[4]), modes = std::atoi(argv[5]), osf = std::atoi(argv[7]);
        double dt = std::atof(argv[6]), bw = std::atof(argv[8]);
        focss::Field f = read_field(argv[2], samples, modes, dt, 0.0);
        std::ofstream out(argv[3], std::ios::binary);
        auto put = [&](const focss::Field& g) {
            out.write(reinterpret_cast<const char*>(g.raw()), (std::streamsize)(g.size() * sizeof(focss::complex_t)));
        };
        focss::Field sampled = focss::rrc_sampling(f, osf);
        put(f.decimate(osf));
        put(f.filter(focss::pulse::rrc(samples / osf, osf, 0.25)));
        put(sampled);
        put(focss::rrc_shaping(sampled, osf));
        put(focss::wdm_select(f, 1, bw));
        put(focss::frequency_shift(f, bw / 3));
        put(focss::bandpass_filter(f, bw / 2));
        put(f.upsample(2));
        put(f.downsample(2));
        put(focss::Field(f).fft_shift());
        put(focss::hard_decision_16qam(sampled));
        focss::Vector<focss::Field> chans(3, f);
        chans[1] = f * 0.5;
        chans[2] = f * 0.25;
        put(focss::wdm_mux(chans, bw));
        return 0;
    }
    if ((cmd == "save-transmission" || cmd == "load-transmission") && argc == 7) {
        int modes = std::atoi(argv[5]), samples = std::atoi(argv[6]);
        if (cmd == "save-transmission") {
            focss::Field tx = read_field(argv[2], samples, modes, 1.0, 0.0);
            focss::Field rx = read_field(argv[3], samples, modes, 1.0, 0.0);
            focss::save_transmission(std::string(argv[4]), tx, rx);
        } else {
            focss::Field tx(focss::Domain{modes, samples, 1.0, 0.0}), rx(focss::Domain{modes, samples, 1.0, 0.0});
            focss::load_transmission(std::string(argv[2]), &tx, &rx);
            write_field(argv[3], tx);
            write_field(argv[4], rx);
        }
        return 0;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
