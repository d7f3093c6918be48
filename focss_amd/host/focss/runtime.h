// Device runtime behind the focss C++ API: one focss_ctx per host thread,
// the arithmetic precision SSFM uses, and the ASE seed.  Every entry aborts
// with the library's message on failure, keeping the reference's void
// signatures (the reference stops on assert).
#ifndef FOCSS_RUNTIME_H_
#define FOCSS_RUNTIME_H_

#include <cstdint>

#include "focss.h"

namespace focss {
focss_ctx* runtime();
void check(int status, const char* what);

// FOCSS_F64 (default) or FOCSS_F32 for SSFM / link programs
void set_precision(int precision);
int precision();

// ASE realisations: Philox key and the stream counter of Amplifier::add_noise
void set_noise_seed(uint64_t seed);
uint64_t noise_seed();
uint32_t next_noise_stream();
}  // namespace focss

#endif  // FOCSS_RUNTIME_H_
