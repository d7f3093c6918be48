// Host-side check driver for the C++ mirror's scoring and pulse functions
// (focss/functions.h, focss/processing/modulation.h): reads
//   int32 modes, samples, cut, n_points; complex<double> tx[modes*samples],
//   rx[modes*samples], points[n_points]
// from argv[1] and prints, one value per line:
//   q2_factor(tx, rx), q2_factor(tx, rx, cut), evm2_factor(tx, rx),
//   bit_error_rate_16qam(tx, rx), bit_error_rate_16qam(tx, rx, cut),
//   symbol_error_rate_16qam(tx, rx, cut), hd_sequence_16qam(rx) codes,
//   demodulate_16qam(points) codes, pulse::rrc(16, 4, 0.25) real parts.
// None of these calls reaches the GPU.
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <vector>

#include "focss/field.h"
#include "focss/functions.h"
#include "focss/processing/modulation.h"

using namespace focss;

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[4];
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    const int modes = hdr[0], samples = hdr[1], cut = hdr[2], n_points = hdr[3];
    Domain d{modes, samples, 1.0, 0.0};
    Field tx(d), rx(d);
    std::vector<complex_t> points(n_points);
    in.read(reinterpret_cast<char*>(tx.raw()), sizeof(complex_t) * tx.size());
    in.read(reinterpret_cast<char*>(rx.raw()), sizeof(complex_t) * rx.size());
    in.read(reinterpret_cast<char*>(points.data()), sizeof(complex_t) * n_points);
    if (!in) return 3;
    std::printf("%.17g\n%.17g\n%.17g\n", q2_factor(tx, rx), q2_factor(tx, rx, cut), evm2_factor(tx, rx));
    std::printf("%.17g\n%.17g\n%.17g\n", bit_error_rate_16qam(tx, rx), bit_error_rate_16qam(tx, rx, cut),
                symbol_error_rate_16qam(tx, rx, cut));
    Vector<int> codes = hd_sequence_16qam(rx);
    for (int i = 0; i < codes.size(); ++i) std::printf("%d\n", codes[i]);
    for (int i = 0; i < n_points; ++i) std::printf("%u\n", demodulate_16qam(points[i]));
    ComplexVector h = pulse::rrc(16, 4, 0.25);
    for (int i = 0; i < h.size(); ++i) std::printf("%.17g\n", h[i].real());
    return 0;
}
