// Two-pass SSFM kernels, double, 2^12 .. 2^15 samples per mode.
#include "registry.h"

namespace focss_hip {

void fill_two_pass_small(Registry<double>& r) {
    reg_two_pass<double, 12>(r);
    reg_two_pass<double, 13>(r);
    reg_two_pass<double, 14>(r);
    reg_two_pass<double, 15>(r);
}

}  // namespace focss_hip
