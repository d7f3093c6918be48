// Memory-only floor of the FP64 two-pass link at 2^18 samples: kernels
// that move the field exactly as the time pass and the spectral pass do
// (same tiles, same row-pair interleaved layout, one read and one write of
// every value) but do no arithmetic, beside a plain copy.  The pass pair of
// focss_link_run cannot run faster than the sum of the two patterns.
//
// Layout between the passes: rows n1 and n1 + 1 (N2 values each) are
// interleaved in runs of W = 4 values (64 B), value (row, col) at
//     (row / 2) 2 N2 + (col / W) 2 W + (row % 2) W + col % W.
// Time pattern: C = 4 columns of all N1 rows per workgroup, 256 threads,
// thread (c, j) owns rows j + 64 r, r < 8, of column c.
// Spectral pattern: one row pair per workgroup, 128 threads, thread (s, j)
// owns values j + 64 r of row s.
//
// Usage: focss_amd/pattern_probe  (prints one line per pattern and batch)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

struct v2d {
    double x, y;
};

constexpr int N1 = 512, N2 = 512, W = 4, C = 4;

__device__ __forceinline__ long long at(int row, int col) {
    return (long long)(row >> 1) * (2 * N2) + (col / W) * (2 * W) + (row & 1) * W + col % W;
}

__global__ __launch_bounds__(256) void k_time_pattern(v2d* f) {
    constexpr int tiles = N2 / C;
    v2d* g = f + (long long)(blockIdx.x / tiles) * N1 * N2;
    const int col = (blockIdx.x % tiles) * C + threadIdx.x % C, j = threadIdx.x / C;
    v2d v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = g[at(j + 64 * r, col)];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r].x += 1.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) g[at(j + 64 * r, col)] = v[r];
}

__global__ __launch_bounds__(128) void k_spec_pattern(v2d* f) {
    constexpr int pairs = N1 / 2;
    v2d* g = f + (long long)(blockIdx.x / pairs) * N1 * N2;
    const int row = (blockIdx.x % pairs) * 2 + threadIdx.x % 2, j = threadIdx.x / 2;
    v2d v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = g[at(row, j + 64 * r)];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r].x += 1.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) g[at(row, j + 64 * r)] = v[r];
}

__global__ __launch_bounds__(256) void k_copy_pattern(const v2d* a, v2d* b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

void check(hipError_t e, int line) {
    if (e != hipSuccess) {
        std::printf("HIP error %s at line %d\n", hipGetErrorString(e), line);
        std::exit(1);
    }
}
#define CHECK(x) check((x), __LINE__)

template <typename F>
double time_us(F launch, int reps) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    launch();
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    for (int i = 0; i < reps; ++i) launch();
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    return 1e3 * ms / reps;
}

}  // namespace

int main() {
    const int reps = 20;
    for (int batch : {64, 128}) {
        const long long n = (long long)batch * N1 * N2;
        v2d *f = nullptr, *h = nullptr;
        CHECK(hipMalloc(&f, n * sizeof(v2d)));
        CHECK(hipMalloc(&h, n * sizeof(v2d)));
        CHECK(hipMemset(f, 0, n * sizeof(v2d)));
        CHECK(hipMemset(h, 0, n * sizeof(v2d)));
        const double bytes = 2.0 * n * sizeof(v2d);
        auto line = [&](const char* what, double us, int passes) {
            std::printf("pattern FP64 2^18 batch %3d (%3lld MiB) %-16s %8.2f us %6.0f GB/s\n", batch,
                        n * (long long)sizeof(v2d) >> 20, what, us, passes * bytes / us * 1e-3);
            std::fflush(stdout);
        };
        const double tt = time_us([&] { k_time_pattern<<<batch * (N2 / C), 256>>>(f); }, reps);
        const double ts = time_us([&] { k_spec_pattern<<<batch * (N1 / 2), 128>>>(f); }, reps);
        const double tc = time_us([&] { k_copy_pattern<<<(unsigned)((n + 255) / 256), 256>>>(f, h, n); }, reps);
        CHECK(hipGetLastError());
        line("time pattern", tt, 1);
        line("spectral pattern", ts, 1);
        line("pair (sum)", tt + ts, 2);
        line("copy", tc, 1);
        CHECK(hipFree(f));
        CHECK(hipFree(h));
    }
    return 0;
}
