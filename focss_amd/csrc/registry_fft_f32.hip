// Standalone DFT kernels, float: whole-in-LDS up to 2^13, two-pass above.
#include "registry.h"

namespace focss_hip {

void fill_fft(FftRegistry<float>& r) {
    reg_fft_whole<float, 1>(r);
    reg_fft_whole<float, 2>(r);
    reg_fft_whole<float, 3>(r);
    reg_fft_whole<float, 4>(r);
    reg_fft_whole<float, 5>(r);
    reg_fft_whole<float, 6>(r);
    reg_fft_whole<float, 7>(r);
    reg_fft_whole<float, 8>(r);
    reg_fft_whole<float, 9>(r);
    reg_fft_whole<float, 10>(r);
    reg_fft_whole<float, 11>(r);
    reg_fft_whole<float, 12>(r);
    reg_fft_whole<float, 13>(r);
    reg_fft_two<float, 14>(r);
    reg_fft_two<float, 15>(r);
    reg_fft_two<float, 16>(r);
    reg_fft_two<float, 17>(r);
    reg_fft_two<float, 18>(r);
    reg_fft_two<float, 19>(r);
    reg_fft_two<float, 20>(r);
}

}  // namespace focss_hip
