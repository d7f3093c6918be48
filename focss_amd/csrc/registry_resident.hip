// LDS-resident SSFM kernels, both precisions.
#include "registry.h"

namespace focss_hip {

template <typename R>
static void fill(Registry<R>& r) {
    reg_resident<R, 64, 1>(r);
    reg_resident<R, 128, 1>(r);
    reg_resident<R, 256, 1>(r);
    reg_resident<R, 512, 1>(r);
    reg_resident<R, 1024, 1>(r);
    reg_resident<R, 2048, 1>(r);
    reg_resident<R, 4096, 1>(r);
    reg_resident<R, 8192, 1>(r);
    reg_resident<R, 64, 2>(r);
    reg_resident<R, 128, 2>(r);
    reg_resident<R, 256, 2>(r);
    reg_resident<R, 512, 2>(r);
    reg_resident<R, 1024, 2>(r);
    reg_resident<R, 2048, 2>(r);
    reg_resident<R, 4096, 2>(r);
}

void fill_resident(Registry<double>& r) { fill(r); }
void fill_resident(Registry<float>& r) { fill(r); }

}  // namespace focss_hip
