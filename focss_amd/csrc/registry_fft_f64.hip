// Standalone DFT kernels, double: whole-in-LDS up to 2^13, two-pass above.
#include "registry.h"

namespace focss_hip {

void fill_fft(FftRegistry<double>& r) {
    reg_fft_whole<double, 1>(r);
    reg_fft_whole<double, 2>(r);
    reg_fft_whole<double, 3>(r);
    reg_fft_whole<double, 4>(r);
    reg_fft_whole<double, 5>(r);
    reg_fft_whole<double, 6>(r);
    reg_fft_whole<double, 7>(r);
    reg_fft_whole<double, 8>(r);
    reg_fft_whole<double, 9>(r);
    reg_fft_whole<double, 10>(r);
    reg_fft_whole<double, 11>(r);
    reg_fft_whole<double, 12>(r);
    reg_fft_whole<double, 13>(r);
    reg_fft_two<double, 14>(r);
    reg_fft_two<double, 15>(r);
    reg_fft_two<double, 16>(r);
    reg_fft_two<double, 17>(r);
    reg_fft_two<double, 18>(r);
    reg_fft_two<double, 19>(r);
    reg_fft_two<double, 20>(r);
}

}  // namespace focss_hip
