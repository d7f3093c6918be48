// Two-pass SSFM kernels, double, 2^16 .. 2^20 samples per mode.
#include "registry.h"

namespace focss_hip {

void fill_two_pass_large(Registry<double>& r) {
    reg_two_pass<double, 16>(r);
    reg_two_pass<double, 17>(r);
    reg_two_pass<double, 18>(r);
    reg_two_pass<double, 19>(r);
    reg_two_pass<double, 20>(r);
}

}  // namespace focss_hip
