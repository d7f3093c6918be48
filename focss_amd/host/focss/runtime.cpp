#include "focss/runtime.h"

#include <cstdio>
#include <cstdlib>

namespace focss {
namespace {
struct Holder {
    focss_ctx* ctx = nullptr;
    ~Holder() {
        if (ctx) focss_ctx_destroy(ctx);
    }
};
thread_local Holder holder;
int g_precision = FOCSS_F64;
uint64_t g_seed = 0xF0C55000ull;
uint32_t g_stream = 0;
thread_local uint64_t g_transfers = 0;
}  // namespace

void count_transfer() { ++g_transfers; }
uint64_t transfer_count() { return g_transfers; }

void check(int status, const char* what) {
    if (status != 0) {
        std::fprintf(stderr, "focss: %s failed (%d): %s\n", what, status, focss_last_error());
        std::abort();
    }
}

focss_ctx* runtime() {
    if (!holder.ctx) check(focss_ctx_create(0, &holder.ctx), "focss_ctx_create");
    return holder.ctx;
}

void set_precision(int p) { g_precision = p; }
int precision() { return g_precision; }
void set_noise_seed(uint64_t seed) {
    g_seed = seed;
    g_stream = 0;
}
uint64_t noise_seed() { return g_seed; }
uint32_t next_noise_stream() { return g_stream++; }
}  // namespace focss
