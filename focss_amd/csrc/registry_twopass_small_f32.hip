// Two-pass SSFM kernels, float, 2^12 .. 2^15 samples per mode.
#include "registry.h"

namespace focss_hip {

void fill_two_pass_small(Registry<float>& r) {
    reg_two_pass<float, 12>(r);
    reg_two_pass<float, 13>(r);
    reg_two_pass<float, 14>(r);
    reg_two_pass<float, 15>(r);
}

}  // namespace focss_hip
