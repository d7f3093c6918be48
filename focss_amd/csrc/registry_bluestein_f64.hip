// Fused Bluestein link kernels, double: Bluestein lengths 2^12 .. 2^20.
#include "registry.h"

namespace focss_hip {

void fill_blu(BluRegistry<double>& r) {
    reg_blu<double, 12>(r);
    reg_blu<double, 13>(r);
    reg_blu<double, 14>(r);
    reg_blu<double, 15>(r);
    reg_blu<double, 16>(r);
    reg_blu<double, 17>(r);
    reg_blu<double, 18>(r);
    reg_blu<double, 19>(r);
    reg_blu<double, 20>(r);
}

}  // namespace focss_hip
