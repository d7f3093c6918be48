// Fused Bluestein link kernels, float: Bluestein lengths 2^12 .. 2^20.
#include "registry.h"

namespace focss_hip {

void fill_blu(BluRegistry<float>& r) {
    reg_blu<float, 12>(r);
    reg_blu<float, 13>(r);
    reg_blu<float, 14>(r);
    reg_blu<float, 15>(r);
    reg_blu<float, 16>(r);
    reg_blu<float, 17>(r);
    reg_blu<float, 18>(r);
    reg_blu<float, 19>(r);
    reg_blu<float, 20>(r);
}

}  // namespace focss_hip
