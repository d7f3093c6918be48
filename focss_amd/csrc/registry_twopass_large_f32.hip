// Two-pass SSFM kernels, float, 2^16 .. 2^20 samples per mode.
#include "registry.h"

namespace focss_hip {

void fill_two_pass_large(Registry<float>& r) {
    reg_two_pass<float, 16>(r);
    reg_two_pass<float, 17>(r);
    reg_two_pass<float, 18>(r);
    reg_two_pass<float, 19>(r);
    reg_two_pass<float, 20>(r);
}

}  // namespace focss_hip
