// Device runtime behind the focss C++ API: one focss_ctx per host thread,
// the arithmetic precision SSFM uses, and the ASE seed.  Every entry aborts
// with the library's message on failure, keeping the reference's void
// signatures (the reference stops on assert).
#ifndef FOCSS_RUNTIME_H_
#define FOCSS_RUNTIME_H_

#include <cstddef>
#include <cstdint>

#include "focss.h"

namespace focss {
focss_ctx* runtime();
void check(int status, const char* what);

// FOCSS_F64 (default) or FOCSS_F32 for SSFM / link programs
void set_precision(int precision);
int precision();

// Host<->device copies of Field data made so far on this thread (each
// upload or download of a field counts one)
void count_transfer();
uint64_t transfer_count();

// ASE realisations: Philox key and the stream counter of Amplifier::add_noise
void set_noise_seed(uint64_t seed);
uint64_t noise_seed();
uint32_t next_noise_stream();

// A device buffer of the runtime's context, filled from host data or left
// uninitialised; freed with the object.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    explicit DeviceBuffer(size_t b) : bytes(b) { check(focss_device_alloc(runtime(), bytes, &ptr), "focss_device_alloc"); }
    DeviceBuffer(const void* host, size_t b) : DeviceBuffer(b) {
        check(focss_memcpy_h2d(runtime(), ptr, host, bytes), "focss_memcpy_h2d");
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (ptr) focss_device_free(runtime(), ptr);
    }
    void download(void* host) const { check(focss_memcpy_d2h(runtime(), host, ptr, bytes), "focss_memcpy_d2h"); }
};
}  // namespace focss

#endif  // FOCSS_RUNTIME_H_
