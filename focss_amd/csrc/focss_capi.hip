// C-ABI of the focss MI355X library (declarations in include/focss.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/focss.h"
#include "registry.h"
#include "triplet_kernels.h"
#include "ldbp_kernels.h"
#include "tmlp_kernels.h"
#include "dsp_kernels.h"
#include "bluestein_kernels.h"

using namespace focss_hip;

// Work buffer of focss_fft_noalloc (the two-pass DFT's intermediate), one
// per device for the life of the process.  Host code never uses this name as
// an address: focss_ctx_create looks the device address up once per device
// (hipGetSymbolAddress) and the entry point passes that address on.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// plain 16-byte copy: the calibration kernel for HBM byte counters
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(FOCSS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Fault injection (focss_inject_fault): the next g_failing_allocs device
// allocations of the library fail as an exhausted device would.
std::atomic<int> g_failing_allocs{0};

template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
    for (int n = g_failing_allocs.load(); n > 0;)
        if (g_failing_allocs.compare_exchange_weak(n, n - 1)) return hipErrorOutOfMemory;
    return hipMalloc(p, bytes);
}

int ilog2_int(int n) {
    int p = 0;
    while ((1 << p) < n) ++p;
    return ((1 << p) == n) ? p : -1;
}

}  // namespace

// ------------------------------------------------------------------
// context
// ------------------------------------------------------------------
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<unsigned char> key;
};

struct BlueTable {
    void* chirp = nullptr;
    void* bhat = nullptr;
    void* bhat_rows = nullptr;  // row order of the fused link kernels (log_m >= kBluMinLog)
    int log_m = 0;
};

struct focss_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void* bank[2] = {nullptr, nullptr};  // [F64, F32]
    void* fine[2] = {nullptr, nullptr};
    // per-call device scratch (grown on demand)
    WaveStatic* d_statics = nullptr;
    WaveState* d_states = nullptr;
    Reduction* d_red = nullptr;
    double* d_partials = nullptr;  // [3][cap][kMaxTimeTiles] per-tile power sums
    int* d_waves = nullptr;
    int* d_done = nullptr;
    int* h_done = nullptr;  // pinned, [2]
    hipEvent_t ev_done[2] = {nullptr, nullptr};
    // LDS-resident length classes run on a second stream beside the two-pass
    // passes (they fill only one workgroup per waveform)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // capture-only streams: the pass-pair graph holds one chain of launches
    // per lane, so that classes on different lanes run side by side
    static constexpr int kLanes = 4;
    int lanes = kLanes;  // FOCSS_GRAPH_LANES=1..4
    hipStream_t lane[kLanes - 1] = {};
    hipEvent_t ev_lane_fork = nullptr;
    hipEvent_t ev_lane_join[kLanes - 1] = {};
    size_t cap_waves = 0;
    int tile_variant[2][2][21] = {};  // [pass: 0 time, 1 spectral][precision][log2 samples]
    GraphCache graph;
    void* scratch = nullptr;  // out-of-place buffer of the two-pass DFT
    size_t scratch_bytes = 0;
    void* noalloc_scratch = nullptr;  // device address of `scratch` on this context's device
    std::map<std::pair<int, int>, BlueTable> blue;  // (length, precision) -> Bluestein tables
    void* blu_work = nullptr;  // standalone Bluestein DFT rows
    size_t blu_work_bytes = 0;
    void* gen_work = nullptr;  // link programs of generic length: rows of the Bluestein convolutions
    size_t gen_work_bytes = 0;
    // padded-convolution kernels of UNIFORM generic-length programs:
    // (length, precision, beta2, s, dw as bits) -> Khat_s in row order
    std::map<std::tuple<int, int, unsigned long long, unsigned long long, unsigned long long>, void*> pad;
    size_t pad_bytes = 0;
    void* pad_tmp = nullptr;  // FP64 staging of one table (N + M values)
    size_t pad_tmp_bytes = 0;
    bool pad_uniform = true;  // FOCSS_BLUESTEIN_UNIFORM=1: UNIFORM programs through Bluestein too
    bool energy_comp = true;  // FOCSS_FP32_ENERGY=0: complex64 programs without the compensated energy
    int split_parts = 2;  // FOCSS_SPLIT_CLASSES=k: chains of a large two-pass class in the graph (1 = one chain)
    void* d_genwave = nullptr;  // [cap] GenWave
    int* d_active = nullptr;    // [cap]
    int graph_pairs = 48;  // multiple of 6 (state parity 2, reduction slots 3)
    long long max_split_steps = 20000000;
    bool use_graph = true;  // FOCSS_NO_GRAPH=1: plain launches (profilers)
    bool natural_layout = false;  // FOCSS_NATURAL_LAYOUT=1: no row-pair interleave between passes
    int n_cu = 256;          // compute units of the device
    void* ldbp_frag = nullptr;  // learned-DBP inference: matrix-core operands of every step
    size_t ldbp_frag_bytes = 0;
    double* ldbp_acc = nullptr;  // learned-DBP training: loss and gradient accumulators
    size_t ldbp_acc_bytes = 0;
    unsigned short* tmlp_w1b = nullptr;  // triplet MLP: bf16 copy of W1 [128][2240]
    double* tmlp_acc = nullptr;          // triplet MLP: loss, then the gradient [TM_NP]
    double* lin_stats = nullptr;         // constant equalisers: the four sums of k_linear_stats
    void* dsp_tmp = nullptr;    // transmitter/receiver DSP work space
    size_t dsp_tmp_bytes = 0;
};

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> address of `scratch`

// The device address of `scratch` on the current device, looked up once per
// device.
int scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it == g_scratch_addr.end()) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        it = g_scratch_addr.emplace(device, p).first;
    }
    *out = it->second;
    return 0;
}

template <typename R>
void host_tables(std::vector<cx<R>>& bank, std::vector<cx<R>>& fine) {
    bank.assign(TW_BANK_SIZE, mk<R>(0, 0));
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int p = 1; p <= TW_MAX_LOG; ++p) {
        int n = 1 << p;
        for (int m = 0; m < n; ++m) {
            long double a = two_pi * (long double)m / (long double)n;
            bank[tw_offset(p) + m] = mk<R>((R)cosl(a), (R)(-sinl(a)));
        }
    }
    fine.assign(1024, mk<R>(0, 0));
    for (int m = 0; m < 1024; ++m) {
        long double a = two_pi * (long double)m / (long double)(1 << 20);
        fine[m] = mk<R>((R)cosl(a), (R)(-sinl(a)));
    }
}

constexpr int kMinSplitWaves = 16;  // two-pass classes of this many waveforms run as two chains of the graph
constexpr int kMaxTimeTiles = 1024;  // time-pass tiles per waveform (N2 <= 1024)
// lengths that are not powers of two (Bluestein DFTs of length 2^20 at most);
// their time op's chunks (kGenTimeChunk samples) also fit kMaxTimeTiles
constexpr int kMaxGenericSamples = 1 << 19;
static_assert(kMaxGenericSamples / kGenTimeChunk <= kMaxTimeTiles, "generic time chunks per waveform");

void free_control(focss_ctx* c) {
    for (void* p : {(void*)c->d_statics, (void*)c->d_states, (void*)c->d_red, (void*)c->d_waves,
                    (void*)c->d_partials, c->d_genwave, (void*)c->d_active})
        if (p) (void)hipFree(p);
    c->d_partials = nullptr;
    c->d_genwave = nullptr;
    c->d_active = nullptr;
    c->d_statics = nullptr;
    c->d_states = nullptr;
    c->d_red = nullptr;
    c->d_waves = nullptr;
    c->cap_waves = 0;
}

// control arrays for n_waves waveforms; on failure the context holds none
// and the next call tries again
int ensure_capacity(focss_ctx* c, size_t n_waves) {
    if (n_waves <= c->cap_waves) return 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamSynchronize(c->side));
    free_control(c);
    const size_t cap = n_waves < 64 ? 64 : n_waves;
    hipError_t e = dev_malloc(&c->d_statics, cap * sizeof(WaveStatic));
    if (e == hipSuccess) e = dev_malloc(&c->d_states, 2 * cap * sizeof(WaveState));
    if (e == hipSuccess) e = dev_malloc(&c->d_red, 3 * cap * sizeof(Reduction));
    if (e == hipSuccess) e = dev_malloc(&c->d_waves, cap * sizeof(int));
    if (e == hipSuccess) e = dev_malloc(&c->d_partials, 3 * cap * kMaxTimeTiles * sizeof(double));
    if (e == hipSuccess) e = dev_malloc(&c->d_genwave, cap * sizeof(GenWave<double>));
    if (e == hipSuccess) e = dev_malloc(&c->d_active, cap * sizeof(int));
    if (e != hipSuccess) {
        free_control(c);
        return fail(FOCSS_E_HIP, std::string("link control arrays: ") + hipGetErrorString(e));
    }
    c->cap_waves = cap;
    return 0;
}

void drop_graph(focss_ctx* c) {
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    if (c->graph.graph) (void)hipGraphDestroy(c->graph.graph);
    c->graph.exec = nullptr;
    c->graph.graph = nullptr;
    c->graph.key.clear();
}

bool resident_ok(int precision, int log_n, int modes) {
    if (precision == FOCSS_F64) return reg<double>().resident[log_n][modes - 1] != nullptr;
    return reg<float>().resident[log_n][modes - 1] != nullptr;
}

bool time_ok(int precision, int log_n, int modes, int variant) {
    if (log_n < 0 || log_n > 20 || variant < 0 || variant >= kTileVariants) return false;
    if (precision == FOCSS_F64) return reg<double>().time_k[log_n][modes - 1][variant] != nullptr;
    return reg<float>().time_k[log_n][modes - 1][variant] != nullptr;
}

bool spec_ok(int precision, int log_n, int variant) {
    if (log_n < 0 || log_n > 20 || variant < 0 || variant >= kTileVariants) return false;
    if (precision == FOCSS_F64) return reg<double>().spec_k[log_n][variant] != nullptr;
    return reg<float>().spec_k[log_n][variant] != nullptr;
}

bool two_pass_ok(int precision, int log_n, int modes, int variant) {
    return time_ok(precision, log_n, modes, variant) && spec_ok(precision, log_n, variant);
}

// the selected variant if it exists for this class, else the first that does
int usable_variant(bool time_pass, int precision, int log_n, int modes, int want) {
    auto ok = [&](int v) { return time_pass ? time_ok(precision, log_n, modes, v) : spec_ok(precision, log_n, v); };
    if (ok(want)) return want;
    for (int v = 0; v < kTileVariants; ++v)
        if (ok(v)) return v;
    return -1;
}

struct LengthClass {
    int log_n, modes, resident;
    int variant;       // time pass tile variant
    int spec_variant;  // spectral pass tile variant
    int layout_g;      // intermediate layout of the two-pass programs (Interleave)
    int first;  // index into the waves array
    int count;
};

template <typename R>
void launch_time(focss_ctx* c, const LengthClass& lc, PassArgs a, Tables<R> tb, hipStream_t s = nullptr) {
    const Registry<R>& r = reg<R>();
    int cols = r.time_cols[lc.log_n][lc.modes - 1][lc.variant];
    int n2 = 1 << (lc.log_n - pass_log_n1(lc.log_n));
    a.tiles_per_wave = n2 / cols;
    a.waves += lc.first;
    a.log_n = lc.log_n;
    a.layout_g = lc.layout_g;
    KernelFn<R> k = r.time_k[lc.log_n][lc.modes - 1][lc.variant];
    const int groups = lc.count * (a.tiles_per_wave / r.time_tpb[lc.log_n][lc.modes - 1][lc.variant]);
    hipLaunchKernelGGL(k, dim3(groups), dim3(r.time_threads[lc.log_n][lc.modes - 1][lc.variant]), 0,
                       s ? s : c->stream, a, tb);
}

template <typename R>
void launch_spec(focss_ctx* c, const LengthClass& lc, PassArgs a, Tables<R> tb, hipStream_t s = nullptr) {
    const Registry<R>& r = reg<R>();
    int rows = r.spec_rows[lc.log_n][lc.spec_variant];
    int n1 = 1 << pass_log_n1(lc.log_n);
    a.tiles_per_wave = lc.modes * n1 / rows;
    a.waves += lc.first;
    a.log_n = lc.log_n;
    a.layout_g = lc.layout_g;
    a.time_tiles = (1 << (lc.log_n - pass_log_n1(lc.log_n))) / r.time_cols[lc.log_n][lc.modes - 1][lc.variant];
    KernelFn<R> k = r.spec_k[lc.log_n][lc.spec_variant];
    const int groups = lc.count * (a.tiles_per_wave / r.spec_tpb[lc.log_n][lc.spec_variant]);
    hipLaunchKernelGGL(k, dim3(groups), dim3(r.spec_threads[lc.log_n][lc.spec_variant]), 0, s ? s : c->stream, a,
                       tb);
}

// natural order back after the programs of an interleaved class end
template <typename R>
void launch_deinterleave(focss_ctx* c, const LengthClass& lc, PassArgs a) {
    const int n1 = 1 << pass_log_n1(lc.log_n);
    const int n2 = 1 << (lc.log_n - pass_log_n1(lc.log_n));
    a.tiles_per_wave = lc.modes * n1 / 2;
    a.waves += lc.first;
    a.log_n = lc.log_n;
    a.layout_g = lc.layout_g;
    hipLaunchKernelGGL(k_deinterleave<R>, dim3(lc.count * a.tiles_per_wave), dim3(256),
                       (size_t)2 * n2 * sizeof(cx<R>), c->stream, a, n2);
}

template <typename R>
void launch_resident(hipStream_t s, const LengthClass& lc, PassArgs a, Tables<R> tb) {
    const Registry<R>& r = reg<R>();
    a.waves += lc.first;
    a.log_n = lc.log_n;
    a.tiles_per_wave = 1;
    hipLaunchKernelGGL(r.resident[lc.log_n][lc.modes - 1], dim3(lc.count), dim3(NT), 0, s, a, tb);
}

int fill_static(const focss_link& L, long long offset, WaveStatic& w) {
    if (L.modes < 1 || L.modes > 2) return fail(FOCSS_E_ARG, "modes must be 1 or 2");
    const int log_n = ilog2_int(L.samples);
    if (!((log_n >= 6 && log_n <= 20) || (L.samples >= 2 && L.samples <= kMaxGenericSamples)))
        return fail(FOCSS_E_ARG, "samples must be a power of two in [2^6, 2^20] or any length in [2, 2^19]");
    if (!(L.dt > 0)) return fail(FOCSS_E_ARG, "dt must be > 0");
    if (L.n_spans < 1) return fail(FOCSS_E_ARG, "n_spans must be >= 1");
    if (!(L.fiber.length > 0)) return fail(FOCSS_E_ARG, "fibre length must be > 0");
    if (L.direction != 1.0 && L.direction != -1.0) return fail(FOCSS_E_ARG, "direction must be +1 or -1");
    const int g = L.grid.grid;
    if (g == FOCSS_UNIFORM) {
        if (L.grid.total_steps < 1) return fail(FOCSS_E_ARG, "UNIFORM grid needs total_steps >= 1");
    } else if (g == FOCSS_LOGARITHMIC) {
        if (L.fiber.alpha == 0) return fail(FOCSS_E_ARG, "LOGARITHMIC grid needs alpha != 0 (ssfm.cpp:76)");
        if (!(L.grid.max_step_size > 0)) return fail(FOCSS_E_ARG, "LOGARITHMIC grid needs max_step_size > 0");
        if (!(L.grid.max_phase_shift > 0)) return fail(FOCSS_E_ARG, "LOGARITHMIC grid needs max_phase_shift > 0");
        if (L.fiber.gamma == 0 || L.fiber.kappa == 0)
            return fail(FOCSS_E_ARG, "LOGARITHMIC grid needs gamma, kappa != 0");
    } else if (g == FOCSS_ADAPTIVE) {
        if (!(L.grid.max_step_size > 0))
            return fail(FOCSS_E_ARG, "ADAPTIVE grid needs max_step_size > 0 (h = 0 never terminates, ssfm.cpp:99)");
        if (!(L.grid.max_phase_shift > 0)) return fail(FOCSS_E_ARG, "ADAPTIVE grid needs max_phase_shift > 0");
    } else {
        return fail(FOCSS_E_ARG, "unknown grid");
    }
    w.alpha = L.fiber.alpha;
    w.beta2 = L.fiber.beta2;
    w.gamma = L.fiber.gamma;
    w.kappa = L.fiber.kappa;
    w.length = L.fiber.length;
    w.kgamma = L.fiber.kappa * L.fiber.gamma;
    w.abs_kgamma = std::fabs(L.fiber.kappa * L.fiber.gamma);
    w.dir = L.direction;
    w.phi_max = L.grid.max_phase_shift;
    w.h_max = L.grid.max_step_size;
    w.pre_scale = L.span_pre_scale;
    w.post_scale = L.span_post_scale;
    w.noise_sigma = L.noise_sigma;
    w.dw = 2.0 * 3.14159265358979323846264338 / (L.dt * L.samples);
    w.seed = L.noise_seed;
    w.offset = offset;
    w.grid = g;
    w.steps = L.grid.total_steps;
    w.n_spans = L.n_spans;
    w.modes = L.modes;
    w.samples = L.samples;
    w.linear_only = (g == FOCSS_UNIFORM && w.kgamma == 0) ? 1 : 0;
    w.sched_mode = 0;
    w.sched_cap = 0;
    w.sched = nullptr;
    w.energy_comp = 0;
    return 0;
}

// ------------------------------------------------------------------
// DFTs of rows (power-of-two lengths), Bluestein tables and the
// standalone DFT of any length
// ------------------------------------------------------------------
template <typename R>
Tables<R> tables_for(focss_ctx* c) {
    Tables<R> tb;
    const int prec = sizeof(R) == 8 ? FOCSS_F64 : FOCSS_F32;
    tb.bank = (const cx<R>*)c->bank[prec];
    tb.fine = (const cx<R>*)c->fine[prec];
    tb.bank64 = (const cx<double>*)c->bank[FOCSS_F64];
    tb.fine64 = (const cx<double>*)c->fine[FOCSS_F64];
    return tb;
}

// a context buffer of at least `need` bytes, grown after the stream drains
int grow(focss_ctx* c, void** buf, size_t* have, size_t need) {
    if (*have >= need) return 0;
    HIP_TRY(hipStreamSynchronize(c->stream));  // earlier work may still use the old buffer
    if (*buf) HIP_TRY(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
    HIP_TRY(dev_malloc(buf, need));
    *have = need;
    return 0;
}

// unscaled-or-scaled DFT (exponent sign `sign`, result x `scale`) of `rows`
// contiguous rows of 2^log_m values; lengths above 2^13 go through `tmp`
// (rows x 2^log_m values)
template <typename R>
void fft_rows_launch(focss_ctx* c, const Tables<R>& tb, int sign, int rows, int log_m, cx<R>* data, cx<R>* tmp,
                     double scale) {
    const FftRegistry<R>& r = fft_reg<R>();
    const int si = sign < 0 ? 0 : 1;
    if (r.whole[log_m][si]) {
        hipLaunchKernelGGL(r.whole[log_m][si], dim3(rows), dim3(NT), 0, c->stream, data, scale, tb);
        return;
    }
    const int n = 1 << log_m;
    const int n1 = 1 << (log_m / 2);
    const int n2 = n / n1;
    const int tc = n2 / r.ncols[log_m];
    const int tr = n1 / r.nrows[log_m];
    hipLaunchKernelGGL(r.cols[log_m][si], dim3(rows * tc), dim3(NT), 0, c->stream, (const cx<R>*)data, tmp, log_m, tc,
                       tb);
    hipLaunchKernelGGL(r.rows[log_m][si], dim3(rows * tr), dim3(NT), 0, c->stream, (const cx<R>*)tmp, data, log_m, tr,
                       scale, tb);
}

int ceil_log2(long long v) {
    int p = 0;
    while ((1LL << p) < v) ++p;
    return p;
}

int fft_enqueue_pow2_f64(focss_ctx* c, int modes, int log_n, void* data);

// The chirp c_j = exp(-i pi j^2 / N) (j^2 reduced mod 2N, in long double)
// and Bhat = DFT_M(b) / M of one length N at precision R (also in the row
// order of the fused link kernels when M >= 2^kBluMinLog), built in FP64
// once per context and kept.
template <typename R>
int blue_table(focss_ctx* c, int n, const cx<R>** chirp, const cx<R>** bhat, int* log_m,
               const cx<R>** bhat_rows = nullptr) {
    const int prec = sizeof(R) == 8 ? FOCSS_F64 : FOCSS_F32;
    auto key = std::make_pair(n, prec);
    auto it = c->blue.find(key);
    if (it == c->blue.end()) {
        if (n < 2 || n > kMaxGenericSamples)
            return fail(FOCSS_E_ARG, "samples must be in [2, 2^19] when not a power of two");
        const int lm = ceil_log2(2LL * n - 1);
        const int M = 1 << lm;
        const long double pi = 3.141592653589793238462643383279502884L;
        std::vector<cx<double>> ch(n), b((size_t)M, mk<double>(0, 0));
        for (long long j = 0; j < n; ++j) {
            const long long q = (j * j) % (2LL * n);
            const long double ang = pi * (long double)q / (long double)n;
            ch[j] = mk<double>((double)cosl(ang), (double)-sinl(ang));
        }
        b[0] = conj(ch[0]);
        for (int j = 1; j < n; ++j) b[j] = b[M - j] = conj(ch[j]);
        BlueTable t;
        t.log_m = lm;
        const bool rows = lm >= kBluMinLog;
        void *d_ch = nullptr, *d_b = nullptr, *d_br = nullptr;  // FP64 tables
        auto drop = [&](hipError_t e) {
            (void)hipStreamSynchronize(c->stream);
            for (void* p : {d_ch, d_b, d_br, t.chirp, t.bhat, t.bhat_rows})
                if (p) (void)hipFree(p);
            return fail(FOCSS_E_HIP, std::string("Bluestein table: ") + hipGetErrorString(e));
        };
        hipError_t e = dev_malloc(&d_ch, (size_t)n * sizeof(cx<double>));
        if (e == hipSuccess) e = dev_malloc(&d_b, (size_t)M * sizeof(cx<double>));
        if (e == hipSuccess && rows) e = dev_malloc(&d_br, (size_t)M * sizeof(cx<double>));
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_ch, ch.data(), (size_t)n * sizeof(cx<double>), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_b, b.data(), (size_t)M * sizeof(cx<double>), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return drop(e);
        if (int rc = fft_enqueue_pow2_f64(c, 1, lm, d_b)) {  // DFT(b) / M
            drop(hipSuccess);
            return rc;
        }
        if (rows)
            hipLaunchKernelGGL(k_blu_row_order<double>, dim3((M + 255) / 256), dim3(256), 0, c->stream,
                               (const cx<double>*)d_b, (cx<double>*)d_br, lm, blu_log_n1(lm));
        if (prec == FOCSS_F32) {
            e = dev_malloc(&t.chirp, (size_t)n * sizeof(cx<float>));
            if (e == hipSuccess) e = dev_malloc(&t.bhat, (size_t)M * sizeof(cx<float>));
            if (e == hipSuccess && rows) e = dev_malloc(&t.bhat_rows, (size_t)M * sizeof(cx<float>));
            if (e != hipSuccess) return drop(e);
            hipLaunchKernelGGL(k_from_c128<float>, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                               (const cx<double>*)d_ch, (cx<float>*)t.chirp, (long long)n);
            hipLaunchKernelGGL(k_from_c128<float>, dim3((M + 255) / 256), dim3(256), 0, c->stream,
                               (const cx<double>*)d_b, (cx<float>*)t.bhat, (long long)M);
            if (rows)
                hipLaunchKernelGGL(k_from_c128<float>, dim3((M + 255) / 256), dim3(256), 0, c->stream,
                                   (const cx<double>*)d_br, (cx<float>*)t.bhat_rows, (long long)M);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return drop(e);
        if (prec == FOCSS_F64) {
            t.chirp = d_ch;
            t.bhat = d_b;
            t.bhat_rows = d_br;
        } else {
            for (void* p : {d_ch, d_b, d_br})
                if (p) (void)hipFree(p);
        }
        it = c->blue.emplace(key, t).first;
    }
    *chirp = (const cx<R>*)it->second.chirp;
    *bhat = (const cx<R>*)it->second.bhat;
    *log_m = it->second.log_m;
    if (bhat_rows) *bhat_rows = (const cx<R>*)it->second.bhat_rows;
    return 0;
}

// stream-ordered DFT of `modes` rows of `samples` (forward x 1/N, inverse
// unscaled).  Power-of-two lengths above 2^13 need a work buffer of the
// field's size: `work` when given (it must hold modes x samples values),
// else the context's buffer, grown on demand.  Other lengths (up to 2^19)
// take the Bluestein path through the context's buffers.
template <typename R>
int fft_enqueue(focss_ctx* c, int sign, int modes, int samples, void* data, void* work = nullptr) {
    if (modes < 1) return fail(FOCSS_E_ARG, "fft: modes must be >= 1");
    const int log_n = ilog2_int(samples);
    const Tables<R> tb = tables_for<R>(c);
    cx<R>* d = (cx<R>*)data;
    if (log_n >= 1 && log_n <= 20) {
        const size_t bytes = (size_t)modes * samples * sizeof(cx<R>);
        if (!work && !fft_reg<R>().whole[log_n][0])
            if (int rc = grow(c, &c->scratch, &c->scratch_bytes, bytes)) return rc;
        fft_rows_launch<R>(c, tb, sign, modes, log_n, d, (cx<R>*)(work ? work : c->scratch),
                           sign < 0 ? 1.0 / samples : 1.0);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (samples < 2 || samples > kMaxGenericSamples)
        return fail(FOCSS_E_ARG, "fft: samples must be a power of two up to 2^20, or any length in [2, 2^19]");
    const cx<R>* chirp;
    const cx<R>* bhat;
    int lm;
    if (int rc = blue_table<R>(c, samples, &chirp, &bhat, &lm)) return rc;
    const size_t mbytes = (size_t)modes * ((size_t)1 << lm) * sizeof(cx<R>);
    if (int rc = grow(c, &c->blu_work, &c->blu_work_bytes, mbytes)) return rc;
    if (int rc = grow(c, &c->scratch, &c->scratch_bytes, mbytes)) return rc;
    cx<R>* wk = (cx<R>*)c->blu_work;
    const int inverse = sign > 0;
    const int chunks = ((1 << lm) + kGenChunk - 1) / kGenChunk;
    hipLaunchKernelGGL(k_blu_in<R>, dim3(modes * chunks), dim3(256), 0, c->stream, (const cx<R>*)d, wk, chirp, samples,
                       lm, inverse);
    fft_rows_launch<R>(c, tb, -1, modes, lm, wk, (cx<R>*)c->scratch, 1.0);
    hipLaunchKernelGGL(k_blu_mul<R>, dim3(modes * chunks), dim3(256), 0, c->stream, wk, bhat, lm);
    fft_rows_launch<R>(c, tb, +1, modes, lm, wk, (cx<R>*)c->scratch, 1.0);
    hipLaunchKernelGGL(k_blu_out<R>, dim3(modes * chunks), dim3(256), 0, c->stream, (const cx<R>*)wk, d, chirp,
                       samples, lm, inverse, inverse ? 1.0 : 1.0 / samples);
    HIP_TRY(hipGetLastError());
    return 0;
}

int fft_enqueue_pow2_f64(focss_ctx* c, int modes, int log_n, void* data) {
    return fft_enqueue<double>(c, -1, modes, 1 << log_n, data);
}

template <typename R>
int fft_impl(focss_ctx* c, int sign, int modes, int samples, void* data, void* work = nullptr) {
    if (int rc = fft_enqueue<R>(c, sign, modes, samples, data, work)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Waveforms whose length is not a power of two (or below 2^6), grouped by
// the power-of-two length 2^log_m of their Bluestein convolutions
struct GenClass {
    int log_m, modes;
    int fused;         // log_m >= kBluMinLog: the four-launch k_blu_* kernels
    int padded;        // fused UNIFORM programs: the two-launch padded convolution
    int first, count;  // into the waves array
    int max_n;         // longest waveform of the class
    int unused = 0;
    long long work_off;  // its rows' first value in gen_work
};

template <typename R>
GenArgs<R> gen_args(focss_ctx* c, const GenClass& gc, const PassArgs& a) {
    GenArgs<R> g;
    g.p = a;
    g.p.waves += gc.first;
    g.gw = (const GenWave<R>*)c->d_genwave;
    g.work = (cx<R>*)c->gen_work + gc.work_off;
    g.active = c->d_active;
    g.log_m = gc.log_m;
    g.modes = gc.modes;
    g.chunks_per_row = ((1 << gc.log_m) + kGenChunk - 1) / kGenChunk;
    return g;
}

template <typename R>
BluArgs<R> blu_args(focss_ctx* c, const GenClass& gc, const PassArgs& a) {
    BluArgs<R> g;
    g.p = a;
    g.p.waves += gc.first;
    g.gw = (const GenWave<R>*)c->d_genwave;
    g.work = (cx<R>*)c->gen_work + gc.work_off;
    g.active = c->d_active;
    g.log_m = gc.log_m;
    g.tiles = 0;
    return g;
}

// the spectral op of counter a.counter: two Bluestein DFTs around the
// dispersion multiply (whole-row DFT kernels below 2^kBluMinLog, else the
// fused row / column kernels), or one padded convolution (one row launch)
template <typename R>
void launch_gen_spec(focss_ctx* c, const GenClass& gc, const PassArgs& a, const Tables<R>& tb,
                     hipStream_t s = nullptr) {
    if (gc.fused) {
        if (!s) s = c->stream;
        const BluRegistry<R>& r = blu_reg<R>();
        BluArgs<R> g = blu_args<R>(c, gc, a);
        g.tiles = gc.modes * r.row_tiles[gc.log_m];
        const dim3 rows_grid(gc.count * g.tiles), rows_block(r.row_threads[gc.log_m]);
        if (gc.padded) {
            hipLaunchKernelGGL(r.rows_pad[gc.log_m], rows_grid, rows_block, 0, s, g, 0, tb);
            return;
        }
        hipLaunchKernelGGL(r.rows[gc.log_m], rows_grid, rows_block, 0, s, g, 0, tb);
        BluArgs<R> gm = g;
        gm.tiles = r.col_tiles[gc.log_m][gc.modes - 1];
        hipLaunchKernelGGL(r.cols[gc.log_m][gc.modes - 1][BLU_MID], dim3(gc.count * gm.tiles),
                           dim3(r.col_threads[gc.log_m][gc.modes - 1]), 0, s, gm, tb);
        hipLaunchKernelGGL(r.rows[gc.log_m], rows_grid, rows_block, 0, s, g, 1, tb);
        return;
    }
    const GenArgs<R> g = gen_args<R>(c, gc, a);
    const int rows = gc.count * gc.modes;
    const dim3 grid(rows * g.chunks_per_row), block(256);
    hipLaunchKernelGGL(k_gen_in<R>, grid, block, 0, c->stream, g);
    fft_rows_launch<R>(c, tb, -1, rows, gc.log_m, g.work, nullptr, 1.0);
    hipLaunchKernelGGL(k_gen_mul<R>, grid, block, 0, c->stream, g);
    fft_rows_launch<R>(c, tb, +1, rows, gc.log_m, g.work, nullptr, 1.0);
    hipLaunchKernelGGL(k_gen_mid<R>, grid, block, 0, c->stream, g);
    fft_rows_launch<R>(c, tb, -1, rows, gc.log_m, g.work, nullptr, 1.0);
    hipLaunchKernelGGL(k_gen_mul<R>, grid, block, 0, c->stream, g);
    fft_rows_launch<R>(c, tb, +1, rows, gc.log_m, g.work, nullptr, 1.0);
    hipLaunchKernelGGL(k_gen_out<R>, grid, block, 0, c->stream, g);
}

// the time op of counter a.counter (entry: the first one)
template <typename R>
void launch_gen_time(focss_ctx* c, const GenClass& gc, const PassArgs& a, const Tables<R>& tb, bool entry,
                     hipStream_t s = nullptr) {
    if (gc.fused) {
        if (!s) s = c->stream;
        const BluRegistry<R>& r = blu_reg<R>();
        BluArgs<R> g = blu_args<R>(c, gc, a);
        g.tiles = r.col_tiles[gc.log_m][gc.modes - 1];
        const int kind = gc.padded ? (entry ? BLU_PAD_ENTRY : BLU_PAD_OUT) : (entry ? BLU_ENTRY : BLU_OUT);
        hipLaunchKernelGGL(r.cols[gc.log_m][gc.modes - 1][kind], dim3(gc.count * g.tiles),
                           dim3(r.col_threads[gc.log_m][gc.modes - 1]), 0, s, g, tb);
        return;
    }
    GenArgs<R> g = gen_args<R>(c, gc, a);
    g.p.tiles_per_wave = (gc.max_n + kGenTimeChunk - 1) / kGenTimeChunk;
    void (*k)(GenArgs<R>) = gc.modes == 1 ? &k_gen_time<R, 1> : &k_gen_time<R, 2>;
    hipLaunchKernelGGL(k, dim3(gc.count * g.p.tiles_per_wave), dim3(256), 0, c->stream, g);
}

// The signed step lengths of the spectral ops of a UNIFORM program (at most
// two: h / 2 at a program's or ASE span's ends, h between), from the device's
// own planner run on the host; false when the program is not UNIFORM, is one
// linear operator (kappa gamma = 0), or applies more lengths
bool uniform_step_lengths(const WaveStatic& w, double kv[2]) {
    if (w.grid != GRID_UNIFORM || w.linear_only) return false;
    WaveState st = initial_state(w);
    int nv = 0;
    const long long limit = (long long)w.n_spans * ((long long)w.steps + 2) + 4;
    for (long long it = 0; it < limit; ++it) {
        const SpectralOp op = plan_spectral(w, st, 0.0, 0.0);  // UNIFORM plans ignore the power
        if (!op.active) {
            if (nv == 1) kv[1] = kv[0];
            return nv > 0;
        }
        int j = 0;
        while (j < nv && kv[j] != op.s) ++j;
        if (j == nv) {
            if (nv == 2) return false;
            kv[nv++] = op.s;
        }
    }
    return false;
}

constexpr size_t kPadCacheBytes = (size_t)8 << 30;  // tables kept across calls, beyond which they are rebuilt

// Khat_s = DFT_M(K_s) / M in row order at precision R for a UNIFORM program
// of length n (K_s from d_s = IDFT_n(D_s), all in FP64), built on the
// context stream once per (length, precision, beta2, s, dw) and kept
template <typename R>
int pad_table(focss_ctx* c, int n, int log_m, double beta2, double s, double dw, const cx<R>** out) {
    const int prec = sizeof(R) == 8 ? FOCSS_F64 : FOCSS_F32;
    const auto key = std::make_tuple(n, prec, double_to_bits(beta2), double_to_bits(s), double_to_bits(dw));
    auto it = c->pad.find(key);
    if (it == c->pad.end()) {
        const int M = 1 << log_m;
        if (int rc = grow(c, &c->pad_tmp, &c->pad_tmp_bytes, ((size_t)n + M) * sizeof(cx<double>))) return rc;
        cx<double>* u = (cx<double>*)c->pad_tmp;
        cx<double>* kv = u + n;
        void* t = nullptr;
        HIP_TRY(dev_malloc(&t, (size_t)M * sizeof(cx<R>)));
        hipLaunchKernelGGL(k_pad_disp, dim3((n + 255) / 256), dim3(256), 0, c->stream, u, n, 0.5 * beta2 * s, dw);
        int rc = fft_enqueue<double>(c, +1, 1, n, u);  // n d
        if (!rc) {
            hipLaunchKernelGGL(k_pad_kernel, dim3(std::min(M / 256, 4096)), dim3(256), 0, c->stream,
                               (const cx<double>*)u, kv, n, log_m);
            rc = fft_enqueue<double>(c, -1, 1, M, kv);  // DFT_M(K) / M
        }
        if (!rc)
            hipLaunchKernelGGL(k_pad_rows<R>, dim3(std::min(M / 256, 4096)), dim3(256), 0, c->stream,
                               (const cx<double>*)kv, (cx<R>*)t, log_m, blu_log_n1(log_m));
        const hipError_t e = hipGetLastError();
        if (rc || e != hipSuccess) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(t);
            return rc ? rc : fail(FOCSS_E_HIP, std::string("padded-convolution table: ") + hipGetErrorString(e));
        }
        c->pad_bytes += (size_t)M * sizeof(cx<R>);
        it = c->pad.emplace(key, t).first;
    }
    *out = (const cx<R>*)it->second;
    return 0;
}

// Bluestein or padded-convolution tables, GenWave records and work rows of
// the generic classes
template <typename R>
int prepare_generic(focss_ctx* c, int n, const focss_link* links, const std::vector<int>& order,
                    std::vector<GenClass>& gen, const std::vector<WaveStatic>& st,
                    const std::vector<std::array<double, 2>>& kvals) {
    if (gen.empty()) return 0;
    if (c->pad_bytes > kPadCacheBytes) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (auto& kv : c->pad) HIP_TRY(hipFree(kv.second));
        c->pad.clear();
        c->pad_bytes = 0;
    }
    std::vector<GenWave<R>> gw(n);
    long long values = 0;
    for (GenClass& gc : gen) {
        gc.work_off = values;
        gc.max_n = 0;
        for (int k = 0; k < gc.count; ++k) {
            const int i = order[gc.first + k];
            GenWave<R>& w = gw[i];
            w.chirp = w.bhat = w.bhat_rows = nullptr;
            w.khat[0] = w.khat[1] = nullptr;
            w.kval[0] = w.kval[1] = 0;
            if (gc.padded) {
                for (int j = 0; j < 2; ++j) {
                    w.kval[j] = kvals[i][j];
                    if (int rc = pad_table<R>(c, links[i].samples, gc.log_m, st[i].beta2, kvals[i][j], st[i].dw,
                                              &w.khat[j]))
                        return rc;
                }
            } else {
                int lm = 0;
                if (int rc = blue_table<R>(c, links[i].samples, &w.chirp, &w.bhat, &lm, &w.bhat_rows)) return rc;
            }
            w.row0 = (long long)k * gc.modes;
            w.n = links[i].samples;
            w.pad = 0;
            gc.max_n = std::max(gc.max_n, w.n);
        }
        values += (long long)gc.count * gc.modes << gc.log_m;
    }
    if (int rc = grow(c, &c->gen_work, &c->gen_work_bytes, (size_t)values * sizeof(cx<R>))) return rc;
    HIP_TRY(hipMemcpyAsync(c->d_genwave, gw.data(), (size_t)n * sizeof(GenWave<R>), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));  // gw is released on return
    return 0;
}

// an ADAPTIVE step schedule to record or replay: sched[i * cap ..] for waveform i
struct Schedule {
    int mode = 0;  // 0 none, FOCSS_SCHEDULE_RECORD, FOCSS_SCHEDULE_REPLAY
    int cap = 0;
    double* sched = nullptr;
};

template <typename R>
int run_links(focss_ctx* c, int precision, int n, const focss_link* links, void* batch,
              const int64_t* offsets, int* total_steps, int profile_pairs = 0, double* prof_ms = nullptr,
              int graph_replays = 0, Schedule schedule = Schedule()) {
    std::vector<WaveStatic> st(n);
    for (int i = 0; i < n; ++i) {
        int rc = fill_static(links[i], offsets[i], st[i]);
        if (rc) return rc;
        st[i].energy_comp = sizeof(R) == 4 && c->energy_comp;
        if (schedule.mode) {
            st[i].sched_mode = schedule.mode;
            st[i].sched_cap = schedule.cap;
            st[i].sched = schedule.sched + (size_t)i * schedule.cap;
        }
    }
    if (int rc = ensure_capacity(c, n)) return rc;
    // group by length class; lengths without a kernel of their own (not a
    // power of two in [2^6, 2^20]) by their Bluestein length, keyed 100 + log_m,
    // and by whether they run the padded convolution
    std::map<std::tuple<int, int, int>, std::vector<int>> groups;
    std::vector<std::array<double, 2>> kvals(n, std::array<double, 2>{0.0, 0.0});
    // the step lengths of a UNIFORM program depend on these fields only
    // (the scales change the gains, not the lengths): one planner run each
    using StepKey = std::tuple<int, int, int, int, unsigned long long, unsigned long long>;
    std::map<StepKey, std::pair<bool, std::array<double, 2>>> step_memo;
    for (int i = 0; i < n; ++i) {
        const int l = ilog2_int(links[i].samples);
        int key = l, padded = 0;
        if (l < 6 || l > 20) {
            const int lm = ceil_log2(2LL * links[i].samples - 1);
            key = 100 + lm;
            if (c->pad_uniform && lm >= kBluMinLog) {
                const WaveStatic& w = st[i];
                const StepKey sk{w.grid * 2 + w.linear_only, w.steps, w.n_spans, w.noise_sigma != 0,
                                 double_to_bits(w.length), double_to_bits(w.dir)};
                auto it = step_memo.find(sk);
                if (it == step_memo.end()) {
                    std::array<double, 2> kv{0.0, 0.0};
                    const bool ok = uniform_step_lengths(w, kv.data());
                    it = step_memo.emplace(sk, std::make_pair(ok, kv)).first;
                }
                padded = it->second.first;
                kvals[i] = it->second.second;
            }
        }
        groups[std::make_tuple(key, links[i].modes, padded)].push_back(i);
    }
    std::vector<int> order;
    std::vector<LengthClass> resident, two_pass;
    std::vector<GenClass> gen;
    for (auto& kv : groups) {
        const int key = std::get<0>(kv.first), modes = std::get<1>(kv.first);
        if (key >= 100) {
            GenClass gc;
            gc.log_m = key - 100;
            gc.modes = modes;
            gc.fused = gc.log_m >= kBluMinLog;
            gc.padded = std::get<2>(kv.first);
            gc.first = (int)order.size();
            gc.count = (int)kv.second.size();
            gc.max_n = 0;
            gc.work_off = 0;
            for (int i : kv.second) order.push_back(i);
            gen.push_back(gc);
            continue;
        }
        LengthClass lc;
        lc.log_n = key;
        lc.modes = modes;
        lc.variant = usable_variant(true, precision, lc.log_n, lc.modes, c->tile_variant[0][precision][lc.log_n]);
        lc.spec_variant = usable_variant(false, precision, lc.log_n, lc.modes, c->tile_variant[1][precision][lc.log_n]);
        lc.first = (int)order.size();
        lc.count = (int)kv.second.size();
        lc.layout_g = 1;
        if (lc.spec_variant >= 0 && !c->natural_layout &&
            reg<R>().spec_rows[lc.log_n][lc.spec_variant] % 2 == 0)
            lc.layout_g = 2;
        for (int i : kv.second) order.push_back(i);
        if (resident_ok(precision, lc.log_n, lc.modes)) {
            lc.resident = 1;
            resident.push_back(lc);
        } else if (lc.variant >= 0 && lc.spec_variant >= 0) {
            lc.resident = 0;
            two_pass.push_back(lc);
        } else {
            return fail(FOCSS_E_UNSUPPORTED, "no kernel for samples=2^" + std::to_string(lc.log_n) +
                                                 " modes=" + std::to_string(lc.modes));
        }
    }
    // in the captured graph a large two-pass class runs as split_parts chains
    // of its waveforms on separate lanes (two by default), so that one part's
    // time pass overlaps another's spectral pass across the kernel boundaries
    // (each waveform's arithmetic is the same: the results do not change)
    const int parts = std::min(c->split_parts, c->lanes);
    if (c->use_graph && parts > 1 && profile_pairs == 0) {
        std::vector<LengthClass> split;
        for (const LengthClass& lc : two_pass) {
            const int k = lc.count >= kMinSplitWaves ? parts : 1;
            for (int p = 0; p < k; ++p) {
                LengthClass h = lc;
                h.first = lc.first + (int)((long long)lc.count * p / k);
                h.count = (int)((long long)lc.count * (p + 1) / k) - (int)((long long)lc.count * p / k);
                split.push_back(h);
            }
        }
        two_pass.swap(split);
    }
    if ((profile_pairs > 0 || graph_replays > 0) && !gen.empty())
        return fail(FOCSS_E_UNSUPPORTED, "pass timing takes power-of-two lengths in [2^6, 2^20] only");
    if (int rc = prepare_generic<R>(c, n, links, order, gen, st, kvals)) return rc;
    std::vector<WaveState> init(2 * (size_t)n);
    for (int i = 0; i < n; ++i) init[i] = initial_state(st[i]);
    std::vector<Reduction> red0(3 * (size_t)n);
    memset(red0.data(), 0, red0.size() * sizeof(Reduction));
    Tables<R> tb;
    tb.bank = (const cx<R>*)c->bank[precision];
    tb.fine = (const cx<R>*)c->fine[precision];
    tb.bank64 = (const cx<double>*)c->bank[FOCSS_F64];
    tb.fine64 = (const cx<double>*)c->fine[FOCSS_F64];
    PassArgs a;
    a.field = batch;
    a.statics = c->d_statics;
    a.states = c->d_states;
    a.red = c->d_red;
    a.done_count = c->d_done;
    a.waves = c->d_waves;
    a.n_waves = n;
    a.counter = 0;
    a.layout_g = 1;
    a.partials = c->d_partials;
    a.partial_stride = kMaxTimeTiles;
    a.time_tiles = 0;
    if (profile_pairs > 0) {
        // measurement mode: the first profile_pairs split steps launched one
        // kernel at a time, each bracketed by events on the context stream
        std::vector<hipEvent_t> ev(2 * (size_t)profile_pairs + 3);
        for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        HIP_TRY(hipMemcpyAsync(c->d_statics, st.data(), n * sizeof(WaveStatic), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_states, init.data(), 2 * n * sizeof(WaveState), hipMemcpyHostToDevice,
                               c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_red, red0.data(), 3 * n * sizeof(Reduction), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_waves, order.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(ev[0], c->stream));
        for (const LengthClass& lc : resident) launch_resident<R>(c->stream, lc, a, tb);
        HIP_TRY(hipEventRecord(ev[1], c->stream));
        for (const LengthClass& lc : two_pass) launch_time<R>(c, lc, a, tb);
        HIP_TRY(hipEventRecord(ev[2], c->stream));
        for (int k = 0; k < profile_pairs && !two_pass.empty(); ++k) {
            PassArgs ak = a;
            ak.counter = k;
            for (const LengthClass& lc : two_pass) launch_spec<R>(c, lc, ak, tb);
            HIP_TRY(hipEventRecord(ev[3 + 2 * k], c->stream));
            ak.counter = k + 1;
            for (const LengthClass& lc : two_pass) launch_time<R>(c, lc, ak, tb);
            HIP_TRY(hipEventRecord(ev[4 + 2 * k], c->stream));
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
        float t = 0;
        double t_res = 0, t_time = 0, t_spec = 0;
        HIP_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
        t_res = t;
        if (!two_pass.empty()) {
            for (int k = 0; k < profile_pairs; ++k) {
                HIP_TRY(hipEventElapsedTime(&t, ev[2 + 2 * k], ev[3 + 2 * k]));
                t_spec += t;
                HIP_TRY(hipEventElapsedTime(&t, ev[3 + 2 * k], ev[4 + 2 * k]));
                t_time += t;
            }
            t_spec /= profile_pairs;
            t_time /= profile_pairs;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        prof_ms[0] = t_time;
        prof_ms[1] = t_spec;
        prof_ms[2] = t_res;
        return 0;
    }
    const bool stepped = !two_pass.empty() || !gen.empty();  // programs driven from the host loop
    if (stepped) {
        // graph of [spectral pass r, time pass r+1] x K, captured before any
        // work is queued; its key is everything baked into the kernel args
        std::vector<unsigned char> key;
        auto put = [&key](const void* p, size_t s) {
            const unsigned char* b = (const unsigned char*)p;
            key.insert(key.end(), b, b + s);
        };
        put(&precision, sizeof(int));
        put(&n, sizeof(int));
        put(&batch, sizeof(void*));
        put(&c->d_statics, sizeof(void*));
        put(&c->d_states, sizeof(void*));
        put(&c->d_red, sizeof(void*));
        put(&c->d_waves, sizeof(void*));
        put(&c->d_partials, sizeof(void*));
        for (const LengthClass& lc : two_pass) put(&lc, sizeof(lc));
        for (const GenClass& gc : gen) put(&gc, sizeof(gc));
        put(&c->gen_work, sizeof(void*));
        put(&c->d_genwave, sizeof(void*));
        put(&c->d_active, sizeof(void*));
        if (c->use_graph && (c->graph.key != key || !c->graph.exec)) {
            drop_graph(c);
            // classes to lanes, the most work first onto the least loaded
            // lane (values moved per split step); the whole-row generic
            // classes, whose DFTs launch on the context stream, stay on lane 0
            std::vector<int> tp_lane(two_pass.size(), 0), gen_lane(gen.size(), 0);
            std::vector<std::pair<double, int>> work;  // (values, unit: two-pass i, or generic ~i)
            double load[focss_ctx::kLanes] = {};
            for (size_t i = 0; i < two_pass.size(); ++i)
                work.push_back({4.0 * two_pass[i].count * two_pass[i].modes * (1 << two_pass[i].log_n), (int)i});
            for (size_t i = 0; i < gen.size(); ++i) {
                const double v = (gen[i].padded ? 5.0 : 10.0) * gen[i].count * gen[i].modes * (1 << gen[i].log_m);
                if (gen[i].fused)
                    work.push_back({v, ~(int)i});
                else
                    load[0] += v;
            }
            std::sort(work.begin(), work.end(), [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
                return x.first > y.first;
            });
            int lanes = 1;
            for (const auto& w : work) {
                int l = 0;
                for (int j = 1; j < std::min(c->lanes, (int)work.size()); ++j)
                    if (load[j] < load[l]) l = j;
                load[l] += w.first;
                lanes = std::max(lanes, l + 1);
                if (w.second >= 0)
                    tp_lane[w.second] = l;
                else
                    gen_lane[~w.second] = l;
            }
            auto lane_stream = [c](int l) { return l == 0 ? c->stream : c->lane[l - 1]; };
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            if (lanes > 1) {
                HIP_TRY(hipEventRecord(c->ev_lane_fork, c->stream));
                for (int l = 1; l < lanes; ++l) HIP_TRY(hipStreamWaitEvent(c->lane[l - 1], c->ev_lane_fork, 0));
            }
            for (int k = 0; k < c->graph_pairs; ++k) {
                PassArgs ak = a;
                ak.counter = k;
                for (size_t i = 0; i < two_pass.size(); ++i)
                    launch_spec<R>(c, two_pass[i], ak, tb, lane_stream(tp_lane[i]));
                for (size_t i = 0; i < gen.size(); ++i)
                    launch_gen_spec<R>(c, gen[i], ak, tb, lane_stream(gen_lane[i]));
                ak.counter = k + 1;
                for (size_t i = 0; i < two_pass.size(); ++i)
                    launch_time<R>(c, two_pass[i], ak, tb, lane_stream(tp_lane[i]));
                for (size_t i = 0; i < gen.size(); ++i)
                    launch_gen_time<R>(c, gen[i], ak, tb, false, lane_stream(gen_lane[i]));
            }
            for (int l = 1; l < lanes; ++l) {
                HIP_TRY(hipEventRecord(c->ev_lane_join[l - 1], c->lane[l - 1]));
                HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_lane_join[l - 1], 0));
            }
            HIP_TRY(hipStreamEndCapture(c->stream, &c->graph.graph));
            HIP_TRY(hipGraphInstantiate(&c->graph.exec, c->graph.graph, nullptr, nullptr, 0));
            c->graph.key = key;
        }
    }
    // a previous call's side-stream work (one that returned early) is done
    // before this call's set-up overwrites the shared state
    HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    HIP_TRY(hipMemcpyAsync(c->d_statics, st.data(), n * sizeof(WaveStatic), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_states, init.data(), 2 * n * sizeof(WaveState), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_red, red0.data(), 3 * n * sizeof(Reduction), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_waves, order.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    *c->h_done = 0;
    HIP_TRY(hipMemcpyAsync(c->d_done, c->h_done, sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (!resident.empty()) {
        // resident classes on the side stream, after the set-up above; each
        // kernel touches only its own waveforms' state, and the done counter
        // is atomic
        HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        for (const LengthClass& lc : resident) launch_resident<R>(c->side, lc, a, tb);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c->ev_join, c->side));
    }
    if (stepped) {
        // entry time pass (counter 0), then replays of the graph
        for (const LengthClass& lc : two_pass) launch_time<R>(c, lc, a, tb);
        for (const GenClass& gc : gen) launch_gen_time<R>(c, gc, a, tb, true);
        HIP_TRY(hipGetLastError());
        if (graph_replays > 0) {
            // measurement mode: `graph_replays` replays of the pass-pair
            // graph between one event pair; the program is left unfinished
            hipEvent_t e0, e1;
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
            HIP_TRY(hipEventRecord(e0, c->stream));
            for (int rep = 0; rep < graph_replays; ++rep) HIP_TRY(hipGraphLaunch(c->graph.exec, c->stream));
            HIP_TRY(hipEventRecord(e1, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (!resident.empty()) HIP_TRY(hipStreamSynchronize(c->side));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            prof_ms[0] = ms / ((double)graph_replays * c->graph_pairs);
            return 0;
        }
        // two replays in flight: replay k+1 is queued before the done
        // counter read after replay k is inspected, so the GPU never waits
        // on the host; surplus passes after completion exit at once
        const int target = n;  // resident waves also count themselves done
        for (long long rep = 0;; ++rep) {
            if (c->use_graph) {
                HIP_TRY(hipGraphLaunch(c->graph.exec, c->stream));
            } else {
                for (int k = 0; k < c->graph_pairs; ++k) {
                    PassArgs ak = a;
                    ak.counter = k;
                    for (const LengthClass& lc : two_pass) launch_spec<R>(c, lc, ak, tb);
                    for (const GenClass& gc : gen) launch_gen_spec<R>(c, gc, ak, tb);
                    ak.counter = k + 1;
                    for (const LengthClass& lc : two_pass) launch_time<R>(c, lc, ak, tb);
                    for (const GenClass& gc : gen) launch_gen_time<R>(c, gc, ak, tb, false);
                }
                HIP_TRY(hipGetLastError());
            }
            HIP_TRY(hipMemcpyAsync(&c->h_done[rep & 1], c->d_done, sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
            HIP_TRY(hipEventRecord(c->ev_done[rep & 1], c->stream));
            if (rep >= 1) {
                HIP_TRY(hipEventSynchronize(c->ev_done[(rep - 1) & 1]));
                if (c->h_done[(rep - 1) & 1] >= target) break;
            }
            if (rep > c->max_split_steps / c->graph_pairs)
                return fail(FOCSS_E_NUMERIC, "link program did not terminate within the split-step limit");
        }
    }
    for (const LengthClass& lc : two_pass)
        if (lc.layout_g == 2) launch_deinterleave<R>(c, lc, a);
    HIP_TRY(hipGetLastError());
    if (!resident.empty()) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    std::vector<WaveState> fin(2 * (size_t)n);
    HIP_TRY(hipMemcpyAsync(fin.data(), c->d_states, 2 * n * sizeof(WaveState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::string faults;
    for (int i = 0; i < n; ++i) {
        // the larger count of the two buffers is the final one
        const WaveState& s0 = fin[i];
        const WaveState& s1 = fin[n + i];
        const WaveState& last = s0.total_steps >= s1.total_steps ? s0 : s1;
        if (total_steps) total_steps[i] = last.total_steps;
        if ((s0.fault || s1.fault) && faults.size() < 200)
            faults += " waveform " + std::to_string(i) + " (span " + std::to_string(last.span) + ", after " +
                      std::to_string(last.total_steps) + " split steps);";
    }
    if (!faults.empty()) return fail(FOCSS_E_NUMERIC, "non-finite field power, program stopped:" + faults);
    return 0;
}

}  // namespace

namespace {

// pack a (2R+1)^2 row-major weight matrix W[(R+m)(2R+1) + R+n] into diamond
// order; m_outer selects (m outer, n inner) or (n outer, m inner)
std::vector<cx<double>> diamond_pack(const double* w, int radius, bool m_outer) {
    std::vector<cx<double>> d;
    const int side = 2 * radius + 1;
    for (int a = -radius; a <= radius; ++a) {
        int b = radius - (a < 0 ? -a : a);
        for (int c = -b; c <= b; ++c) {
            int m = m_outer ? a : c, n = m_outer ? c : a;
            const double* e = w + 2 * ((radius + m) * side + (radius + n));
            d.push_back(mk<double>(e[0], e[1]));
        }
    }
    return d;
}

template <typename R>
int upload_weights(focss_ctx* c, const std::vector<cx<double>>& d, void** dev) {
    std::vector<cx<R>> h(d.size());
    for (size_t i = 0; i < d.size(); ++i) h[i] = mk<R>((R)d[i].x, (R)d[i].y);
    HIP_TRY(dev_malloc(dev, std::max<size_t>(1, h.size()) * sizeof(cx<R>)));
    HIP_TRY(hipMemcpyAsync(*dev, h.data(), h.size() * sizeof(cx<R>), hipMemcpyHostToDevice, c->stream));
    return 0;
}

template <typename R>
int sino_cubic_impl(focss_ctx* c, int radius, const double* weights, int modes, int samples, const void* in,
                    void* out) {
    std::vector<cx<double>> d = diamond_pack(weights, radius, true);
    void* w = nullptr;
    if (int rc = upload_weights<R>(c, d, &w)) return rc;
    int blocks = (samples + TRIP_NT - 1) / TRIP_NT;
    size_t lds = (size_t)modes * (TRIP_NT + 2 * radius) * sizeof(cx<R>);
    if (lds > 64 * 1024) {
        (void)hipFree(w);
        return fail(FOCSS_E_UNSUPPORTED, "sino_cubic: radius too large for the LDS window");
    }
    hipLaunchKernelGGL(k_sino_cubic<R>, dim3(blocks), dim3(TRIP_NT), lds, c->stream, (const cx<R>*)in, (cx<R>*)out,
                       (const cx<R>*)w, radius, modes, samples);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(w);
    if (e != hipSuccess) return fail(FOCSS_E_HIP, std::string("sino_cubic: ") + hipGetErrorString(e));
    return 0;
}

// Hermitian positive-definite solve A x = b by Cholesky A = L L^H
int cholesky_solve(std::vector<cx<double>>& A, int n, std::vector<cx<double>>& b) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j].x;
        for (int k = 0; k < j; ++k) d -= norm2(A[(size_t)j * n + k]);
        if (!(d > 0)) return fail(FOCSS_E_NUMERIC, "sino_train: normal matrix not positive definite");
        double l = std::sqrt(d);
        A[(size_t)j * n + j] = mk<double>(l, 0);
        for (int i = j + 1; i < n; ++i) {
            cx<double> v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v = v - A[(size_t)i * n + k] * conj(A[(size_t)j * n + k]);
            A[(size_t)i * n + j] = v * (1.0 / l);
        }
    }
    for (int i = 0; i < n; ++i) {
        cx<double> v = b[i];
        for (int k = 0; k < i; ++k) v = v - A[(size_t)i * n + k] * b[k];
        b[i] = v * (1.0 / A[(size_t)i * n + i].x);
    }
    for (int i = n - 1; i >= 0; --i) {
        cx<double> v = b[i];
        for (int k = i + 1; k < n; ++k) v = v - conj(A[(size_t)k * n + i]) * b[k];
        b[i] = v * (1.0 / A[(size_t)i * n + i].x);
    }
    return 0;
}

// the folded weights (one per class {(m,n), (n,m), (-m,-n), (-n,-m)}, m in
// [-R/2, R/2], n in [|m|, R - |m|]) written into all four places of the
// (2R+1)^2 matrix, interleaved doubles              sino.cpp:124-153
void unfold_symmetric(const std::vector<cx<double>>& b, int* index, int radius, double* out) {
    const int side = 2 * radius + 1;
    for (int i = 0; i < 2 * side * side; ++i) out[i] = 0;
    auto put = [&](int m, int n, cx<double> v) {
        double* e = out + 2 * ((size_t)(radius + m) * side + (radius + n));
        e[0] = v.x;
        e[1] = v.y;
    };
    const int mb = radius / 2;
    for (int m = -mb; m <= mb; ++m) {
        const int am = m < 0 ? -m : m;
        for (int n = am; n <= radius - am; ++n) {
            cx<double> v = b[(*index)++];
            put(m, n, v);
            if (m == 0 && n == 0) continue;
            if (m == n) {
                put(-m, -n, v);
            } else if (m == -n) {
                put(n, m, v);
            } else {
                put(n, m, v);
                put(-m, -n, v);
                put(-n, -m, v);
            }
        }
    }
}

// The least squares of SINO::train / WdmPPE::train (sino.cpp:105-122,
// wdm_sino.cpp:84-100) from device features F (column-major rows x nf) and
// responses y: F^H F, column sums and F^H y on the device, then on the host
// z-scoring (std with 1/rows, arma stddev norm_type 1), centring, the normal
// equations by Cholesky, de-standardisation and the intercept.
int zscore_normal_solve(focss_ctx* c, const cx<double>* F, const cx<double>* y, long long rows, int nf,
                        std::vector<cx<double>>* weights, cx<double>* intercept) {
    cx<double>* G = nullptr;
    cx<double>* sums = nullptr;
    cx<double>* fy = nullptr;
    hipError_t e = dev_malloc(&G, (size_t)nf * nf * sizeof(cx<double>));
    if (e == hipSuccess) e = dev_malloc(&sums, (size_t)(nf + 1) * sizeof(cx<double>));
    if (e == hipSuccess) e = dev_malloc(&fy, (size_t)nf * sizeof(cx<double>));
    std::vector<cx<double>> hG((size_t)nf * nf), hs(nf + 1), hfy(nf);
    if (e == hipSuccess) e = hipMemsetAsync(G, 0, (size_t)nf * nf * sizeof(cx<double>), c->stream);
    if (e == hipSuccess) {
        // upper 32 x 32 tiles, split over row slices of at least 8 chunks so
        // that about 1024 workgroups run
        const int side = (nf + GM_T - 1) / GM_T;
        const int pairs = side * (side + 1) / 2;
        const long long max_split = std::max<long long>(1, rows / (8 * GM_KC));
        const int split = (int)std::min<long long>(std::max(1, 1024 / pairs), max_split);
        hipLaunchKernelGGL(k_gram_mfma, dim3(pairs, split), dim3(256), 0, c->stream, F, G, rows, nf, side);
        hipLaunchKernelGGL(k_col_sums, dim3(nf + 1), dim3(256), 0, c->stream, F, y, rows, nf, sums, fy);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(hG.data(), G, hG.size() * sizeof(cx<double>), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hs.data(), sums, hs.size() * sizeof(cx<double>), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hfy.data(), fy, hfy.size() * sizeof(cx<double>), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(G);
    (void)hipFree(sums);
    (void)hipFree(fy);
    if (e != hipSuccess) return fail(FOCSS_E_HIP, std::string("triplet least squares: ") + hipGetErrorString(e));
    // the lower triangle from the upper tiles (G is Hermitian)
    for (int i = 1; i < nf; ++i)
        for (int j = 0; j < i; ++j) hG[(size_t)i * nf + j] = conj(hG[(size_t)j * nf + i]);
    const double inv = 1.0 / (double)rows;
    std::vector<cx<double>> mu(nf);
    std::vector<double> sd(nf);
    for (int j = 0; j < nf; ++j) {
        mu[j] = hs[j] * inv;
        double var = hG[(size_t)j * nf + j].x * inv - norm2(mu[j]);
        sd[j] = std::sqrt(var > 0 ? var : 0.0);
    }
    const cx<double> ymean = hs[nf] * inv;
    std::vector<cx<double>> A((size_t)nf * nf), b(nf);
    for (int i = 0; i < nf; ++i) {
        for (int j = 0; j < nf; ++j) {
            cx<double> v = hG[(size_t)i * nf + j] - conj(mu[i]) * mu[j] * (double)rows;
            A[(size_t)i * nf + j] = v * (1.0 / (sd[i] * sd[j]));
        }
        b[i] = (hfy[i] - conj(mu[i]) * ymean * (double)rows) * (1.0 / sd[i]);
    }
    if (int rc = cholesky_solve(A, nf, b)) return rc;
    cx<double> ic = ymean;
    for (int j = 0; j < nf; ++j) {
        b[j] = b[j] * (1.0 / sd[j]);
        ic = ic - b[j] * mu[j];
    }
    *intercept = ic;
    weights->swap(b);
    return 0;
}

}  // namespace

namespace {

// ---------------- transmitter / receiver DSP ----------------
inline unsigned grid_for(long long n) { return (unsigned)((n + 255) / 256); }

int dsp_workspace(focss_ctx* c, size_t bytes, char** out) {
    if (c->dsp_tmp_bytes < bytes) {
        HIP_TRY(hipStreamSynchronize(c->stream));  // earlier work may still use the old buffer
        if (c->dsp_tmp) (void)hipFree(c->dsp_tmp);
        c->dsp_tmp = nullptr;
        c->dsp_tmp_bytes = 0;
        HIP_TRY(dev_malloc(&c->dsp_tmp, bytes));
        c->dsp_tmp_bytes = bytes;
    }
    *out = (char*)c->dsp_tmp;
    return 0;
}

int dsp_check(int modes, int samples) {
    if (modes < 1) return fail(FOCSS_E_ARG, "dsp: modes must be >= 1");
    int l = ilog2_int(samples);
    if (l < 1 || l > 20) return fail(FOCSS_E_ARG, "dsp: samples must be a power of two in [2, 2^20]");
    return 0;
}

template <typename R>
int frequency_shift_impl(focss_ctx* c, int modes, int samples, double dt, double frequency, void* data) {
    const long long total = (long long)modes * samples;
    hipLaunchKernelGGL(k_frequency_shift<R>, dim3(grid_for(total)), dim3(256), 0, c->stream, (cx<R>*)data, total,
                       samples, 2.0 * M_PI * frequency * dt);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename R>
int bandpass_impl(focss_ctx* c, int modes, int samples, double dt, double frequency, void* data) {
    const int cut = (int)(frequency * (samples * dt)) + 1;  // modulation.cpp:266-277
    const long long total = (long long)modes * samples;
    if (int rc = fft_enqueue<R>(c, -1, modes, samples, data)) return rc;
    hipLaunchKernelGGL(k_band_zero<R>, dim3(grid_for(total)), dim3(256), 0, c->stream, (cx<R>*)data, total, samples,
                       cut);
    HIP_TRY(hipGetLastError());
    return fft_enqueue<R>(c, +1, modes, samples, data);
}

// Field::filter_inplace (field.cpp:358-372): the impulse response (taps,
// centred at sample 0, odd lengths drop the last tap) is zero-padded to the
// field length and transformed with the 1/N convention of focss::fft, then
// multiplied into the field's spectrum.  spec_ws: samples complex of R.
template <typename R>
int filter_impl(focss_ctx* c, int modes, int samples, int taps, const cx<double>* h_host, void* data, cx<R>* spec_ws,
                cx<double>* upload_ws) {
    std::vector<cx<double>> zp((size_t)samples, mk<double>(0, 0));
    const int half = taps / 2, shift = samples - half;
    for (int i = 0; i < half; ++i) {
        zp[i] = h_host[i + half];
        zp[shift + i] = h_host[i];
    }
    HIP_TRY(hipMemcpyAsync(upload_ws, zp.data(), zp.size() * sizeof(cx<double>), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_from_c128<R>, dim3(grid_for(samples)), dim3(256), 0, c->stream, upload_ws, spec_ws,
                       (long long)samples);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));  // the host copy of zp goes out of scope
    if (int rc = fft_enqueue<R>(c, -1, 1, samples, spec_ws)) return rc;
    if (int rc = fft_enqueue<R>(c, -1, modes, samples, data)) return rc;
    const long long total = (long long)modes * samples;
    hipLaunchKernelGGL(k_mul_spectrum<R>, dim3(grid_for(total)), dim3(256), 0, c->stream, (cx<R>*)data,
                       (const cx<R>*)spec_ws, total, samples);
    HIP_TRY(hipGetLastError());
    return fft_enqueue<R>(c, +1, modes, samples, data);
}

// pulse::rrc (modulation.cpp:47-74): root-raised-cosine taps over the frame,
// normalised to power `size`
std::vector<cx<double>> rrc_pulse_host(int n_symbols, int osf, double roll_off) {
    const int size = n_symbols * osf;
    std::vector<cx<double>> p((size_t)size);
    double power = 0;
    for (int i = 0; i < size; ++i) {
        const double t = (double)i / osf - n_symbols / 2.0;
        double v;
        if (t == 0) {
            v = 1 - roll_off + 4 * roll_off / M_PI;
        } else if (std::fabs(t) == 1 / roll_off / 4) {
            v = roll_off / std::sqrt(2.0) *
                ((1 + 2 / M_PI) * std::sin(M_PI / roll_off / 4) + (1 - 2 / M_PI) * std::cos(M_PI / roll_off / 4));
        } else {
            v = (4 * roll_off * t * std::cos(M_PI * t * (1 + roll_off)) + std::sin(M_PI * t * (1 - roll_off))) /
                (1 - 16 * roll_off * roll_off * t * t) / (M_PI * t);
        }
        p[i] = mk<double>(v, 0);
        power += v * v;
    }
    for (auto& x : p) x = x * (size / power);
    return p;
}

template <typename R>
int decimate_impl(focss_ctx* c, int modes, int samples, int factor, const void* in, void* out) {
    const int ds = samples / factor;
    const size_t bytes = (size_t)modes * samples * sizeof(cx<R>);
    char* ws;
    if (int rc = dsp_workspace(c, bytes, &ws)) return rc;
    HIP_TRY(hipMemcpyAsync(ws, in, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (int rc = fft_enqueue<R>(c, -1, modes, samples, ws)) return rc;
    hipLaunchKernelGGL(k_decimate_gather<R>, dim3(grid_for((long long)modes * ds)), dim3(256), 0, c->stream,
                       (const cx<R>*)ws, (cx<R>*)out, modes, samples, ds, samples / factor / 2);
    HIP_TRY(hipGetLastError());
    return fft_enqueue<R>(c, +1, modes, ds, out);
}

template <typename R>
int rrc_shaping_impl(focss_ctx* c, int modes, int n_symbols, const void* symbols, int osf, double roll_off,
                     double gain, void* out) {
    const int samples = n_symbols * osf;
    const long long total = (long long)modes * samples;
    char* ws;
    if (int rc = dsp_workspace(c, (size_t)samples * (sizeof(cx<double>) + sizeof(cx<R>)), &ws)) return rc;
    HIP_TRY(hipMemsetAsync(out, 0, (size_t)total * sizeof(cx<R>), c->stream));
    hipLaunchKernelGGL(k_upsample<R>, dim3(grid_for((long long)modes * n_symbols)), dim3(256), 0, c->stream,
                       (const cx<R>*)symbols, (cx<R>*)out, (long long)modes * n_symbols, osf);
    HIP_TRY(hipGetLastError());
    const std::vector<cx<double>> rrc = rrc_pulse_host(n_symbols, osf, roll_off);
    if (int rc = filter_impl<R>(c, modes, samples, samples, rrc.data(), out,
                                (cx<R>*)(ws + (size_t)samples * sizeof(cx<double>)), (cx<double>*)ws))
        return rc;
    hipLaunchKernelGGL(k_scale<R>, dim3(grid_for(total)), dim3(256), 0, c->stream, (cx<R>*)out, total,
                       std::sqrt(gain) * osf);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename R>
int rrc_sampling_impl(focss_ctx* c, int modes, int samples, const void* signal, int osf, double roll_off, double gain,
                      void* out) {
    const int n_symbols = samples / osf;
    const size_t work = (size_t)modes * samples * sizeof(cx<R>);
    char* ws;
    if (int rc = dsp_workspace(c, work + (size_t)samples * (sizeof(cx<double>) + sizeof(cx<R>)), &ws)) return rc;
    cx<R>* tmp = (cx<R>*)ws;
    HIP_TRY(hipMemcpyAsync(tmp, signal, work, hipMemcpyDeviceToDevice, c->stream));
    const std::vector<cx<double>> rrc = rrc_pulse_host(n_symbols, osf, roll_off);
    if (int rc = filter_impl<R>(c, modes, samples, samples, rrc.data(), tmp,
                                (cx<R>*)(ws + work + (size_t)samples * sizeof(cx<double>)),
                                (cx<double>*)(ws + work)))
        return rc;
    hipLaunchKernelGGL(k_downsample<R>, dim3(grid_for((long long)modes * n_symbols)), dim3(256), 0, c->stream,
                       (const cx<R>*)tmp, (cx<R>*)out, (long long)modes * n_symbols, osf, std::sqrt(gain));
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename R>
int wdm_select_impl(focss_ctx* c, int modes, int samples, double dt, int channel_index, double bandwidth,
                           const void* in, void* out) {
    HIP_TRY(hipMemcpyAsync(out, in, (size_t)modes * samples * sizeof(cx<R>), hipMemcpyDeviceToDevice, c->stream));
    if (int rc = frequency_shift_impl<R>(c, modes, samples, dt, -bandwidth * channel_index, out)) return rc;
    return bandpass_impl<R>(c, modes, samples, dt, 0.5 * bandwidth, out);
}

template <typename R>
int filter_entry(focss_ctx* c, int modes, int samples, int taps, const double* h, void* data) {
    char* ws;
    if (int rc = dsp_workspace(c, (size_t)samples * (sizeof(cx<double>) + sizeof(cx<R>)), &ws)) return rc;
    return filter_impl<R>(c, modes, samples, taps, (const cx<double>*)h, data,
                          (cx<R>*)(ws + (size_t)samples * sizeof(cx<double>)), (cx<double>*)ws);
}

// records <-> field: the tile is 2^11 values, (2^11 >> te_log2) samples by
// 2^te_log2 envelopes, te_log2 = min(ceil(log2 envelopes), 5)
template <typename R>
int envelope_transpose(focss_ctx* c, bool to_field, int envelopes, int samples, const void* in, void* out) {
    int te_log2 = 0;
    while ((1 << te_log2) < envelopes && te_log2 < 5) ++te_log2;
    const int ts = IO_TILE >> te_log2;
    const dim3 grid((unsigned)((samples + ts - 1) / ts), (unsigned)((envelopes + (1 << te_log2) - 1) >> te_log2));
    if (to_field)
        hipLaunchKernelGGL((k_envelope_transpose<R, true>), grid, dim3(IO_NT), 0, c->stream, in, out, samples,
                           envelopes, te_log2);
    else
        hipLaunchKernelGGL((k_envelope_transpose<R, false>), grid, dim3(IO_NT), 0, c->stream, in, out, samples,
                           envelopes, te_log2);
    HIP_TRY(hipGetLastError());
    return 0;
}

int records_check(focss_ctx* c, int envelopes, int samples, const void* in, void* out, const char* what) {
    if (!c || !in || !out) return fail(FOCSS_E_ARG, std::string(what) + ": null argument");
    if (envelopes < 1 || envelopes > (1 << 20) || samples < 1)
        return fail(FOCSS_E_ARG, std::string(what) + ": envelopes must be in [1, 2^20] and samples >= 1");
    if (in == out) return fail(FOCSS_E_ARG, std::string(what) + ": in and out must not alias");
    return 0;
}

}  // namespace

namespace {

using LdbpKernel = void (*)(LdbpArgs);
using LdbpFwdStepKernel = void (*)(LdbpArgs, int, double*);
using LdbpStepKernel = void (*)(LdbpArgs, int, const cx<float>*, cx<float>*);

// the kernels of one (taps, kerr taps, modes) model: the whole-waveform
// forward (inference), and the per-step forward and backward of training
struct LdbpSet {
    LdbpKernel fwd;
    LdbpFwdStepKernel fstep;
    LdbpStepKernel bwd;
};

template <int K, int KP, int M>
LdbpSet ldbp_set() {
    return {&k_ldbp_fwd<K, KP, M>, &k_ldbp_fwd_step<K, KP, M>, &k_ldbp_bwd_step<K, KP, M>};
}

template <int K, int M>
bool ldbp_kernels_km(int ktaps, LdbpSet* set) {
    switch (ktaps) {
        case 1: *set = ldbp_set<K, 1, M>(); return true;
        case 5: *set = ldbp_set<K, 5, M>(); return true;
        case 9: *set = ldbp_set<K, 9, M>(); return true;
        case 15: *set = ldbp_set<K, 15, M>(); return true;
        default: return false;
    }
}

template <int K>
bool ldbp_kernels_k(int ktaps, int modes, LdbpSet* set) {
    if (modes == 1) return ldbp_kernels_km<K, 1>(ktaps, set);
    if (modes == 2) return ldbp_kernels_km<K, 2>(ktaps, set);
    return false;
}

bool ldbp_kernels(int taps, int ktaps, int modes, LdbpSet* set) {
    if (taps == 15) return ldbp_kernels_k<15>(ktaps, modes, set);
    if (taps == 23) return ldbp_kernels_k<23>(ktaps, modes, set);
    if (taps == 31) return ldbp_kernels_k<31>(ktaps, modes, set);
    if (taps == 47) return ldbp_kernels_k<47>(ktaps, modes, set);
    if (taps == 63) return ldbp_kernels_k<63>(ktaps, modes, set);
    if (taps == 5) return ldbp_kernels_k<5>(ktaps, modes, set);
    return false;
}

int ldbp_check(int steps, int taps, int ktaps, int modes, int samples, int batch) {
    if (modes != 1 && modes != 2) return fail(FOCSS_E_ARG, "learned DBP: modes must be 1 or 2");
    LdbpSet set;
    if (!ldbp_kernels(taps, ktaps, modes, &set))
        return fail(FOCSS_E_UNSUPPORTED, "learned DBP: taps must be 5, 15, 23, 31, 47 or 63 and kerr_taps 1, 5, 9 or 15");
    if (steps < 1 || batch < 1) return fail(FOCSS_E_ARG, "learned DBP: steps and batch must be >= 1");
    int l = ilog2_int(samples);
    if (l < 0 || samples < LDBP_MIN_N || (long long)samples * modes > LDBP_MAX_N)
        return fail(FOCSS_E_ARG, "learned DBP: samples must be a power of two in [1024, 16384 / modes]");
    return 0;
}

// forward: one workgroup of N / 16 threads per waveform, every mode in LDS
// in the p16 layout, M (N + N/16) complex64
int ldbp_launch_fwd(focss_ctx* c, LdbpKernel k, const LdbpArgs& a, int modes) {
    hipLaunchKernelGGL(k, dim3(a.B), dim3(a.N / LDBP_R), (size_t)modes * (a.N + a.N / 16) * sizeof(cx<float>),
                       c->stream, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

// training forward: one launch per step, one workgroup per LDBP_CH-sample
// chunk of each waveform; the loss shares of the chunks added in order
int ldbp_launch_fwd_steps(focss_ctx* c, LdbpFwdStepKernel k, const LdbpArgs& a, int modes, double* lslot) {
    const int groups = a.B * (a.N / LDBP_CH);
    for (int j = 0; j < a.J; ++j) {
        hipLaunchKernelGGL(k, dim3(groups), dim3(ldbp_bwd_threads(modes)), 0, c->stream, a, j, lslot);
        HIP_TRY(hipGetLastError());
    }
    if (a.loss && a.target) {
        hipLaunchKernelGGL(k_ldbp_loss_reduce, dim3(1), dim3(256), 0, c->stream, (const double*)lslot, groups, a.loss);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// backward: one launch per step, last step first, one workgroup per
// LDBP_CH-sample chunk of each waveform; G_y alternates between gy0 and gy1
int ldbp_launch_bwd(focss_ctx* c, LdbpStepKernel k, const LdbpArgs& a, int modes, cx<float>* gy0, cx<float>* gy1) {
    const int groups = a.B * (a.N / LDBP_CH);
    for (int j = a.J - 1; j >= 0; --j) {
        cx<float>* out = (j & 1) ? gy1 : gy0;
        const cx<float>* in = (j & 1) ? gy0 : gy1;
        hipLaunchKernelGGL(k, dim3(groups), dim3(ldbp_bwd_threads(modes)), 0, c->stream, a, j, in, out);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ldbp_grad_reduce, dim3(a.J * (2 * a.K + a.KP)), dim3(256), 0, c->stream,
                       (const double*)a.part, groups, a.J, a.K, a.KP, a.grad);
    HIP_TRY(hipGetLastError());
    return 0;
}

using LdbpInferKernel = void (*)(LdbpArgs, const ldbp_h8*);
using LdbpAfragKernel = void (*)(const cx<float>*, int, ldbp_h8*);

template <int K, int M>
bool ldbp_infer_km(int ktaps, LdbpInferKernel* k) {
    switch (ktaps) {
        case 1: *k = &k_ldbp_infer<K, 1, M>; return true;
        case 5: *k = &k_ldbp_infer<K, 5, M>; return true;
        case 9: *k = &k_ldbp_infer<K, 9, M>; return true;
        case 15: *k = &k_ldbp_infer<K, 15, M>; return true;
        default: return false;
    }
}

template <int K>
bool ldbp_infer_k(int ktaps, int modes, LdbpInferKernel* k, LdbpAfragKernel* f, int* nks) {
    *f = &k_ldbp_afrag<K>;
    *nks = InferShape<K>::NKS;
    if (modes == 1) return ldbp_infer_km<K, 1>(ktaps, k);
    if (modes == 2) return ldbp_infer_km<K, 2>(ktaps, k);
    return false;
}

}  // namespace

extern "C" {

const char* focss_last_error(void) { return g_err.c_str(); }

const char* focss_build_info(void) {
    return "focss-mi355x: gfx950 HIP kernels; resident LDS SSFM (N*modes <= 8192 c128 / 16384 c64), "
           "two-pass four-step SSFM (2^12..2^20); FP64 and FP32 (FP64 phase generation)";
}

}  // extern "C"
namespace {
int ctx_init(focss_ctx* c);
}
extern "C" {

int focss_ctx_create(int device, focss_ctx** out) {
    if (!out) return fail(FOCSS_E_ARG, "null out");
    focss_ctx* c = new focss_ctx();
    c->device = device;
    const char* ng = getenv("FOCSS_NO_GRAPH");
    c->use_graph = !(ng && ng[0] == '1');
    const char* nl = getenv("FOCSS_NATURAL_LAYOUT");
    c->natural_layout = nl && nl[0] == '1';
    const char* gl = getenv("FOCSS_GRAPH_LANES");
    if (gl && atoi(gl) >= 1) c->lanes = std::min(atoi(gl), (int)focss_ctx::kLanes);
    const char* bu = getenv("FOCSS_BLUESTEIN_UNIFORM");
    c->pad_uniform = !(bu && bu[0] == '1');
    const char* fe = getenv("FOCSS_FP32_ENERGY");
    c->energy_comp = !(fe && fe[0] == '0');
    const char* sc = getenv("FOCSS_SPLIT_CLASSES");
    if (sc && sc[0]) c->split_parts = std::max(1, std::min(atoi(sc), (int)focss_ctx::kLanes));
    for (int p = 0; p <= 20; ++p) {
        // measured defaults at 2^18 samples: time pass 2048 elements per
        // tile; spectral pass 1024, two tiles per workgroup for complex64
        c->tile_variant[0][FOCSS_F64][p] = 1;
        c->tile_variant[0][FOCSS_F32][p] = 1;
        c->tile_variant[1][FOCSS_F64][p] = 2;
        c->tile_variant[1][FOCSS_F32][p] = 5;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        return fail(FOCSS_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    if (int rc = ctx_init(c)) {
        const std::string msg = g_err;
        focss_ctx_destroy(c);
        return fail(rc, msg);
    }
    *out = c;
    return 0;
}

}  // extern "C"

namespace {
// device resources of a new context; on failure focss_ctx_destroy releases
// what was made
int ctx_init(focss_ctx* c) {
    const int device = c->device;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    if (int rc = scratch_address(device, &c->noalloc_scratch)) return rc;
    std::vector<cx<double>> b64, f64;
    std::vector<cx<float>> b32, f32;
    host_tables<double>(b64, f64);
    host_tables<float>(b32, f32);
    HIP_TRY(dev_malloc(&c->bank[0], b64.size() * sizeof(cx<double>)));
    HIP_TRY(dev_malloc(&c->fine[0], f64.size() * sizeof(cx<double>)));
    HIP_TRY(dev_malloc(&c->bank[1], b32.size() * sizeof(cx<float>)));
    HIP_TRY(dev_malloc(&c->fine[1], f32.size() * sizeof(cx<float>)));
    HIP_TRY(hipMemcpy(c->bank[0], b64.data(), b64.size() * sizeof(cx<double>), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->fine[0], f64.data(), f64.size() * sizeof(cx<double>), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->bank[1], b32.data(), b32.size() * sizeof(cx<float>), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->fine[1], f32.data(), f32.size() * sizeof(cx<float>), hipMemcpyHostToDevice));
    HIP_TRY(dev_malloc(&c->d_done, sizeof(int)));
    HIP_TRY(hipHostMalloc(&c->h_done, 2 * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_done[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_done[1], hipEventDisableTiming));
    HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_lane_fork, hipEventDisableTiming));
    for (int l = 0; l < focss_ctx::kLanes - 1; ++l) {
        HIP_TRY(hipStreamCreateWithFlags(&c->lane[l], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_lane_join[l], hipEventDisableTiming));
    }
    return 0;
}
}  // namespace

extern "C" {

int focss_ctx_destroy(focss_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    drop_graph(c);
    for (int i = 0; i < 2; ++i) {
        if (c->bank[i]) (void)hipFree(c->bank[i]);
        if (c->fine[i]) (void)hipFree(c->fine[i]);
    }
    if (c->ldbp_frag) (void)hipFree(c->ldbp_frag);
    if (c->ldbp_acc) (void)hipFree(c->ldbp_acc);
    if (c->tmlp_w1b) (void)hipFree(c->tmlp_w1b);
    if (c->tmlp_acc) (void)hipFree(c->tmlp_acc);
    if (c->lin_stats) (void)hipFree(c->lin_stats);
    if (c->dsp_tmp) (void)hipFree(c->dsp_tmp);
    free_control(c);
    if (c->d_done) (void)hipFree(c->d_done);
    if (c->scratch) (void)hipFree(c->scratch);
    for (auto& kv : c->blue)
        for (void* p : {kv.second.chirp, kv.second.bhat, kv.second.bhat_rows})
            if (p) (void)hipFree(p);
    c->blue.clear();
    for (auto& kv : c->pad) (void)hipFree(kv.second);
    c->pad.clear();
    for (void* p : {c->blu_work, c->gen_work, c->pad_tmp})
        if (p) (void)hipFree(p);
    if (c->h_done) (void)hipHostFree(c->h_done);
    for (int i = 0; i < 2; ++i)
        if (c->ev_done[i]) (void)hipEventDestroy(c->ev_done[i]);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->ev_lane_fork) (void)hipEventDestroy(c->ev_lane_fork);
    for (int l = 0; l < focss_ctx::kLanes - 1; ++l) {
        if (c->ev_lane_join[l]) (void)hipEventDestroy(c->ev_lane_join[l]);
        if (c->lane[l]) (void)hipStreamDestroy(c->lane[l]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int focss_device_alloc(focss_ctx* c, size_t bytes, void** dptr) {
    if (!c || !dptr) return fail(FOCSS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(dev_malloc(dptr, bytes));
    return 0;
}

int focss_device_free(focss_ctx* c, void* dptr) {
    if (!c) return fail(FOCSS_E_ARG, "null ctx");
    HIP_TRY(hipFree(dptr));
    return 0;
}

int focss_memcpy_h2d(focss_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(FOCSS_E_ARG, "null ctx");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int focss_memcpy_d2h(focss_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(FOCSS_E_ARG, "null ctx");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int focss_synchronize(focss_ctx* c) {
    if (!c) return fail(FOCSS_E_ARG, "null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamSynchronize(c->side));
    return 0;
}

int focss_link_run(focss_ctx* c, int precision, int n_waves, const focss_link* links, void* batch,
                   const int64_t* offsets, int* total_steps) {
    if (!c || !links || !batch || !offsets || n_waves < 1) return fail(FOCSS_E_ARG, "null argument");
    if (precision != FOCSS_F64 && precision != FOCSS_F32) return fail(FOCSS_E_ARG, "bad precision");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64) return run_links<double>(c, precision, n_waves, links, batch, offsets, total_steps);
    return run_links<float>(c, precision, n_waves, links, batch, offsets, total_steps);
}

int focss_link_run_schedule(focss_ctx* c, int precision, int n_waves, const focss_link* links, void* batch,
                            const int64_t* offsets, int* total_steps, int mode, double* schedule, int capacity) {
    if (!c || !links || !batch || !offsets || n_waves < 1 || !schedule) return fail(FOCSS_E_ARG, "null argument");
    if (precision != FOCSS_F64 && precision != FOCSS_F32) return fail(FOCSS_E_ARG, "bad precision");
    if (mode != FOCSS_SCHEDULE_RECORD && mode != FOCSS_SCHEDULE_REPLAY)
        return fail(FOCSS_E_ARG, "schedule mode must be FOCSS_SCHEDULE_RECORD or FOCSS_SCHEDULE_REPLAY");
    if (capacity < 1) return fail(FOCSS_E_ARG, "schedule capacity must be >= 1");
    HIP_TRY(hipSetDevice(c->device));
    Schedule sc;
    sc.mode = mode;
    sc.cap = capacity;
    sc.sched = schedule;
    if (precision == FOCSS_F64)
        return run_links<double>(c, precision, n_waves, links, batch, offsets, total_steps, 0, nullptr, 0, sc);
    return run_links<float>(c, precision, n_waves, links, batch, offsets, total_steps, 0, nullptr, 0, sc);
}

int focss_link_run_host(focss_ctx* c, int precision, const focss_link* link, double* field, int* total_steps) {
    if (!c || !link || !field) return fail(FOCSS_E_ARG, "null argument");
    size_t n = (size_t)link->modes * (size_t)link->samples;
    size_t esz = precision == FOCSS_F64 ? 16 : 8;
    void* d = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(dev_malloc(&d, n * esz));
    std::vector<float> tmp;
    if (precision == FOCSS_F64) {
        HIP_TRY(hipMemcpyAsync(d, field, n * esz, hipMemcpyHostToDevice, c->stream));
    } else {
        tmp.resize(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)field[i];
        HIP_TRY(hipMemcpyAsync(d, tmp.data(), n * esz, hipMemcpyHostToDevice, c->stream));
    }
    int64_t off = 0;
    int rc = focss_link_run(c, precision, 1, link, d, &off, total_steps);
    if (rc == 0) {
        if (precision == FOCSS_F64) {
            HIP_TRY(hipMemcpyAsync(field, d, n * esz, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
        } else {
            HIP_TRY(hipMemcpyAsync(tmp.data(), d, n * esz, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            for (size_t i = 0; i < 2 * n; ++i) field[i] = tmp[i];
        }
    }
    drop_graph(c);  // the graph captured d, which is freed now
    (void)hipFree(d);
    return rc;
}

int focss_link_profile(focss_ctx* c, int precision, int n_waves, const focss_link* links, void* batch,
                       const int64_t* offsets, int n_pairs, double* avg_ms) {
    if (!c || !links || !batch || !offsets || !avg_ms || n_waves < 1 || n_pairs < 1)
        return fail(FOCSS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64)
        return run_links<double>(c, precision, n_waves, links, batch, offsets, nullptr, n_pairs, avg_ms);
    if (precision == FOCSS_F32)
        return run_links<float>(c, precision, n_waves, links, batch, offsets, nullptr, n_pairs, avg_ms);
    return fail(FOCSS_E_ARG, "bad precision");
}

int focss_link_profile_graph(focss_ctx* c, int precision, int n_waves, const focss_link* links, void* batch,
                             const int64_t* offsets, int n_replays, double* ms_per_pair) {
    if (!c || !links || !batch || !offsets || !ms_per_pair || n_waves < 1 || n_replays < 1)
        return fail(FOCSS_E_ARG, "null argument");
    if (!c->use_graph) return fail(FOCSS_E_UNSUPPORTED, "graph replay timing with FOCSS_NO_GRAPH set");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64)
        return run_links<double>(c, precision, n_waves, links, batch, offsets, nullptr, 0, ms_per_pair, n_replays);
    if (precision == FOCSS_F32)
        return run_links<float>(c, precision, n_waves, links, batch, offsets, nullptr, 0, ms_per_pair, n_replays);
    return fail(FOCSS_E_ARG, "bad precision");
}

int focss_device_copy(focss_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src || (bytes % 16) != 0) return fail(FOCSS_E_ARG, "copy needs 16-byte multiples");
    size_t n = bytes / 16;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, c->stream, (const uint4*)src, (uint4*)dst, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

void* focss_ctx_stream(focss_ctx* c) { return c ? (void*)c->stream : nullptr; }

int focss_link_validate(const focss_link* link) {
    if (!link) return fail(FOCSS_E_ARG, "null link");
    WaveStatic w;
    return fill_static(*link, 0, w);
}

int focss_fft(focss_ctx* c, int precision, int sign, int modes, int samples, void* data) {
    if (!c || !data) return fail(FOCSS_E_ARG, "null argument");
    if (sign != -1 && sign != 1) return fail(FOCSS_E_ARG, "fft: sign must be -1 or +1");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64) return fft_impl<double>(c, sign, modes, samples, data);
    if (precision == FOCSS_F32) return fft_impl<float>(c, sign, modes, samples, data);
    return fail(FOCSS_E_ARG, "bad precision");
}

int focss_inject_fault(int kind, int count) {
    if (kind != FOCSS_FAULT_ALLOC || count < 0) return fail(FOCSS_E_ARG, "inject_fault: unknown kind or count < 0");
    g_failing_allocs.store(count);
    return 0;
}

int focss_fft_noalloc(focss_ctx* c, int precision, int sign, int modes, int samples, void* data) {
    if (!c || !data) return fail(FOCSS_E_ARG, "null argument");
    if (sign != -1 && sign != 1) return fail(FOCSS_E_ARG, "fft: sign must be -1 or +1");
    if (precision != FOCSS_F64 && precision != FOCSS_F32) return fail(FOCSS_E_ARG, "bad precision");
    if (modes < 1 || samples < 1) return fail(FOCSS_E_ARG, "fft: modes and samples must be >= 1");
    if (ilog2_int(samples) < 1)
        return fail(FOCSS_E_ARG, "focss_fft_noalloc: samples must be a power of two in [2, 2^20] (other lengths "
                                 "build Bluestein tables on the device: use focss_fft)");
    const size_t bytes = (size_t)modes * (size_t)samples * (precision == FOCSS_F64 ? 16 : 8);
    if (bytes > kNoAllocScratchBytes)
        return fail(FOCSS_E_ARG, "focss_fft_noalloc: the field is " + std::to_string(bytes) +
                                     " bytes, above its " + std::to_string(kNoAllocScratchBytes) +
                                     "-byte work buffer (use focss_fft)");
    if (!c->noalloc_scratch) return fail(FOCSS_E_HIP, "focss_fft_noalloc: no work buffer on this device");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64) return fft_impl<double>(c, sign, modes, samples, data, c->noalloc_scratch);
    return fft_impl<float>(c, sign, modes, samples, data, c->noalloc_scratch);
}

int focss_add_noise(focss_ctx* c, int precision, int modes, int samples, double sigma, uint64_t seed,
                    uint32_t stream, void* data) {
    if (!c || !data || modes < 1 || samples < 1) return fail(FOCSS_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(c->device));
    long long n = (long long)modes * samples;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    if (precision == FOCSS_F64)
        hipLaunchKernelGGL(k_add_noise<double>, dim3(blocks), dim3(256), 0, c->stream, (cx<double>*)data, n, sigma,
                           (unsigned long long)seed, stream);
    else if (precision == FOCSS_F32)
        hipLaunchKernelGGL(k_add_noise<float>, dim3(blocks), dim3(256), 0, c->stream, (cx<float>*)data, n, sigma,
                           (unsigned long long)seed, stream);
    else
        return fail(FOCSS_E_ARG, "bad precision");
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int focss_field_scale(focss_ctx* c, int precision, int modes, int samples, double re, double im, void* data) {
    if (!c || !data || modes < 1 || samples < 1) return fail(FOCSS_E_ARG, "field_scale: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    const long long n = (long long)modes * samples;
    const unsigned blocks = std::min<unsigned>(grid_for(n), 8192u);
    if (precision == FOCSS_F64)
        hipLaunchKernelGGL(k_mul_scalar<double>, dim3(blocks), dim3(256), 0, c->stream, (cx<double>*)data, n, re, im);
    else if (precision == FOCSS_F32)
        hipLaunchKernelGGL(k_mul_scalar<float>, dim3(blocks), dim3(256), 0, c->stream, (cx<float>*)data, n, re, im);
    else
        return fail(FOCSS_E_ARG, "bad precision");
    HIP_TRY(hipGetLastError());
    return 0;
}

int focss_sino_cubic(focss_ctx* c, int precision, int radius, const double* weights, int modes, int samples,
                     const void* in, void* out) {
    if (!c || !weights || !in || !out || radius < 0 || modes < 1 || samples <= 2 * radius)
        return fail(FOCSS_E_ARG, "sino_cubic: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    if (precision == FOCSS_F64) return sino_cubic_impl<double>(c, radius, weights, modes, samples, in, out);
    if (precision == FOCSS_F32) return sino_cubic_impl<float>(c, radius, weights, modes, samples, in, out);
    return fail(FOCSS_E_ARG, "bad precision");
}

int focss_sino_train(focss_ctx* c, int radius, int type, const double* linear_weight_in, int modes, int samples,
                     const void* tx, const void* rx, double* weights, double* intercept, double* linear_weight) {
    if (!c || !tx || !rx || !weights || !intercept || !linear_weight || radius < 0 || type < 0 || type > 2 ||
        modes < 1 || samples <= 2 * radius)
        return fail(FOCSS_E_ARG, "sino_train: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    const int nf = radius * (radius + 2) / 2 + 1 + (type == 0 ? 1 : 0);
    const long long rows = (long long)modes * (samples - 2 * radius);
    if (rows < nf) return fail(FOCSS_E_ARG, "sino_train: fewer rows than features");
    cx<double>* F = nullptr;
    cx<double>* y = nullptr;
    HIP_TRY(dev_malloc(&F, (size_t)rows * nf * sizeof(cx<double>)));
    if (dev_malloc(&y, (size_t)rows * sizeof(cx<double>)) != hipSuccess) {
        (void)hipFree(F);
        return fail(FOCSS_E_HIP, "sino_train: out of device memory");
    }
    cx<double> lw = mk<double>(linear_weight_in ? linear_weight_in[0] : 0.0, linear_weight_in ? linear_weight_in[1] : 0.0);
    int fblocks = (int)std::min<long long>((rows + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sino_features, dim3(fblocks), dim3(256), 0, c->stream, (const cx<double>*)tx,
                       (const cx<double>*)rx, F, y, radius, type, lw, modes, samples);
    std::vector<cx<double>> b;
    cx<double> ic;
    int rc = zscore_normal_solve(c, F, y, rows, nf, &b, &ic);
    (void)hipFree(F);
    (void)hipFree(y);
    if (rc) return rc;
    // unfold                                           sino.cpp:124-153
    intercept[0] = ic.x;
    intercept[1] = ic.y;
    int index = 0;
    unfold_symmetric(b, &index, radius, weights);
    if (type == 0) {
        linear_weight[0] = b[index].x;
        linear_weight[1] = b[index].y;
    } else if (type == 2) {
        linear_weight[0] = 0;
        linear_weight[1] = 0;
    } else {
        linear_weight[0] = lw.x;
        linear_weight[1] = lw.y;
    }
    return 0;
}

int focss_wdm_ppe_train(focss_ctx* c, int c_radius, int d_radius, int samples, const void* tx6, const void* rx6,
                        double* C, double* D, double* intercept, double* linear_weight) {
    const int r = std::max(c_radius, d_radius);
    if (!c || !tx6 || !rx6 || !intercept || !linear_weight || c_radius < -1 || d_radius < -1 ||
        (c_radius >= 0 && !C) || (d_radius >= 0 && !D) || samples <= 2 * std::max(r, 0))
        return fail(FOCSS_E_ARG, "wdm_ppe_train: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    const int n_spm = c_radius < 0 ? 0 : c_radius * (c_radius + 2) / 2 + 1;
    const int n_xpm = d_radius < 0 ? 0 : 2 * d_radius * d_radius + 2 * d_radius + 1;
    const int nf = n_spm + n_xpm + 1;
    const long long rows = 2LL * (samples - 2 * std::max(r, 0));
    if (rows < nf) return fail(FOCSS_E_ARG, "wdm_ppe_train: fewer rows than features");
    cx<double>* F = nullptr;
    cx<double>* y = nullptr;
    HIP_TRY(dev_malloc(&F, (size_t)rows * nf * sizeof(cx<double>)));
    if (dev_malloc(&y, (size_t)rows * sizeof(cx<double>)) != hipSuccess) {
        (void)hipFree(F);
        return fail(FOCSS_E_HIP, "wdm_ppe_train: out of device memory");
    }
    int fblocks = (int)std::min<long long>((rows + 255) / 256, 4096);
    hipLaunchKernelGGL(k_wdmppe_features, dim3(fblocks), dim3(256), 0, c->stream, (const cx<double>*)tx6,
                       (const cx<double>*)rx6, F, y, c_radius, d_radius, samples);
    std::vector<cx<double>> b;
    cx<double> ic;
    int rc = zscore_normal_solve(c, F, y, rows, nf, &b, &ic);
    (void)hipFree(F);
    (void)hipFree(y);
    if (rc) return rc;
    // unfold C (the SINO fold), D (the diamond, m outer), linear weight last
    //                                                   wdm_sino.cpp:102-137
    intercept[0] = ic.x;
    intercept[1] = ic.y;
    int index = 0;
    if (c_radius >= 0) unfold_symmetric(b, &index, c_radius, C);
    if (d_radius >= 0) {
        const int side = 2 * d_radius + 1;
        for (int i = 0; i < 2 * side * side; ++i) D[i] = 0;
        for (int m = -d_radius; m <= d_radius; ++m) {
            const int bd = d_radius - (m < 0 ? -m : m);
            for (int n = -bd; n <= bd; ++n) {
                double* e = D + 2 * ((size_t)(d_radius + m) * side + (d_radius + n));
                e[0] = b[index].x;
                e[1] = b[index].y;
                ++index;
            }
        }
    }
    linear_weight[0] = b[index].x;
    linear_weight[1] = b[index].y;
    return 0;
}

int focss_wdm_ppe(focss_ctx* c, int precision, int c_radius, const double* C, int d_radius, const double* D,
                  int samples, const void* a6, void* spm, void* xpm) {
    if (!c || !a6 || !spm || !xpm || samples < 1 || (c_radius >= 0 && !C) || (d_radius >= 0 && !D))
        return fail(FOCSS_E_ARG, "wdm_ppe: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    std::vector<cx<double>> cd = c_radius >= 0 ? diamond_pack(C, c_radius, false) : std::vector<cx<double>>();
    std::vector<cx<double>> dd = d_radius >= 0 ? diamond_pack(D, d_radius, true) : std::vector<cx<double>>();
    void *cw = nullptr, *dw = nullptr;
    int blocks = (samples + TRIP_NT - 1) / TRIP_NT;
    if (precision == FOCSS_F64) {
        if (int rc = upload_weights<double>(c, cd, &cw)) return rc;
        if (int rc = upload_weights<double>(c, dd, &dw)) return rc;
        hipLaunchKernelGGL(k_wdm_ppe<double>, dim3(blocks), dim3(TRIP_NT), 0, c->stream, (const cx<double>*)a6,
                           (const cx<double>*)cw, c_radius, (const cx<double>*)dw, d_radius, samples, (cx<double>*)spm,
                           (cx<double>*)xpm);
    } else if (precision == FOCSS_F32) {
        if (int rc = upload_weights<float>(c, cd, &cw)) return rc;
        if (int rc = upload_weights<float>(c, dd, &dw)) return rc;
        hipLaunchKernelGGL(k_wdm_ppe<float>, dim3(blocks), dim3(TRIP_NT), 0, c->stream, (const cx<float>*)a6,
                           (const cx<float>*)cw, c_radius, (const cx<float>*)dw, d_radius, samples, (cx<float>*)spm,
                           (cx<float>*)xpm);
    } else {
        return fail(FOCSS_E_ARG, "bad precision");
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(cw);
    (void)hipFree(dw);
    if (e != hipSuccess) return fail(FOCSS_E_HIP, std::string("wdm_ppe: ") + hipGetErrorString(e));
    return 0;
}

// activations, final output, two G_y buffers, then the per-workgroup
// gradient sums of every backward step (taps up to 63, kerr taps up to 15)
// and the per-workgroup loss shares
static size_t ldbp_part_doubles(int steps, int samples, int batch) {
    return (size_t)(steps * (2 * 63 + 15) + 1) * batch * (samples / LDBP_CH);
}

size_t focss_ldbp_workspace_bytes(int steps, int modes, int samples, int batch) {
    size_t acts = 2ull * (size_t)batch * steps * modes * samples * sizeof(cx<float>);
    size_t out = (size_t)batch * modes * samples * sizeof(cx<float>);
    return acts + 3 * out + ldbp_part_doubles(steps, samples, batch) * sizeof(double) + 256;
}

int focss_ldbp_forward(focss_ctx* c, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                       const float* params, const void* in, void* out, const void* target, double* loss) {
    if (!c || !params || !in || (!out && !target) || (target && !loss))
        return fail(FOCSS_E_ARG, "learned DBP: null argument");
    if (int rc = ldbp_check(steps, taps, kerr_taps, modes, samples, batch)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    LdbpSet set;
    ldbp_kernels(taps, kerr_taps, modes, &set);
    double* dloss = nullptr;
    if (target) {
        HIP_TRY(dev_malloc(&dloss, sizeof(double)));
        HIP_TRY(hipMemsetAsync(dloss, 0, sizeof(double), c->stream));
    }
    LdbpArgs a = {};
    a.in = (const cx<float>*)in;
    a.out = (cx<float>*)out;
    a.target = (const cx<float>*)target;
    a.fir = (const cx<float>*)params;
    a.kerr = params + 2 * steps * taps;
    a.loss = dloss;
    a.J = steps;
    a.K = taps;
    a.KP = kerr_taps;
    a.N = samples;
    a.B = batch;
    int rc = ldbp_launch_fwd(c, set.fwd, a, modes);
    if (dloss) {
        hipError_t e = hipSuccess;
        if (rc == 0) e = hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(dloss);
        if (e != hipSuccess) return fail(FOCSS_E_HIP, std::string("learned DBP loss: ") + hipGetErrorString(e));
    }
    return rc;
}

int focss_ldbp_infer(focss_ctx* c, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                     const float* params, const void* in, void* out) {
    if (!c || !params || !in || !out) return fail(FOCSS_E_ARG, "learned DBP: null argument");
    if (int rc = ldbp_check(steps, taps, kerr_taps, modes, samples, batch)) return rc;
    LdbpInferKernel k = nullptr;
    LdbpAfragKernel f = nullptr;
    int nks = 0;
    bool ok = false;
    switch (taps) {
        case 5: ok = ldbp_infer_k<5>(kerr_taps, modes, &k, &f, &nks); break;
        case 15: ok = ldbp_infer_k<15>(kerr_taps, modes, &k, &f, &nks); break;
        case 23: ok = ldbp_infer_k<23>(kerr_taps, modes, &k, &f, &nks); break;
        case 31: ok = ldbp_infer_k<31>(kerr_taps, modes, &k, &f, &nks); break;
        case 47: ok = ldbp_infer_k<47>(kerr_taps, modes, &k, &f, &nks); break;
        case 63: ok = ldbp_infer_k<63>(kerr_taps, modes, &k, &f, &nks); break;
        default: break;
    }
    if (!ok) return fail(FOCSS_E_UNSUPPORTED, "learned DBP: no inference kernel for these taps");
    HIP_TRY(hipSetDevice(c->device));
    // A operands of every step, in the context's scratch (grown on demand)
    const size_t frag_bytes = (size_t)steps * nks * 64 * sizeof(ldbp_h8);
    if (c->ldbp_frag_bytes < frag_bytes) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->ldbp_frag) (void)hipFree(c->ldbp_frag);
        c->ldbp_frag = nullptr;
        c->ldbp_frag_bytes = 0;
        HIP_TRY(dev_malloc(&c->ldbp_frag, frag_bytes));
        c->ldbp_frag_bytes = frag_bytes;
    }
    ldbp_h8* afrag = (ldbp_h8*)c->ldbp_frag;
    const int n_frag = steps * nks * 64;
    hipLaunchKernelGGL(f, dim3((n_frag + 255) / 256), dim3(256), 0, c->stream, (const cx<float>*)params, steps, afrag);
    HIP_TRY(hipGetLastError());
    LdbpArgs a = {};
    a.in = (const cx<float>*)in;
    a.out = (cx<float>*)out;
    a.fir = (const cx<float>*)params;
    a.kerr = params + 2 * steps * taps;
    a.J = steps;
    a.K = taps;
    a.KP = kerr_taps;
    a.N = samples;
    a.B = batch;
    const size_t lds = (size_t)modes * samples * (sizeof(cx_h) + sizeof(float));
    hipLaunchKernelGGL(k, dim3(batch), dim3(LI_NT), lds, c->stream, a, (const ldbp_h8*)afrag);
    HIP_TRY(hipGetLastError());
    return 0;
}

int focss_ldbp_gradients(focss_ctx* c, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                         const float* params, const void* in, const void* target, void* workspace,
                         size_t workspace_bytes, double* grad, double* loss) {
    if (!c || !params || !in || !target || !workspace || !grad) return fail(FOCSS_E_ARG, "learned DBP: null argument");
    if (int rc = ldbp_check(steps, taps, kerr_taps, modes, samples, batch)) return rc;
    if (workspace_bytes < focss_ldbp_workspace_bytes(steps, modes, samples, batch))
        return fail(FOCSS_E_ARG, "learned DBP: workspace too small");
    HIP_TRY(hipSetDevice(c->device));
    LdbpSet set;
    ldbp_kernels(taps, kerr_taps, modes, &set);
    size_t act = (size_t)batch * steps * modes * samples;
    cx<float>* xs = (cx<float>*)workspace;
    cx<float>* ys = xs + act;
    cx<float>* out = ys + act;
    cx<float>* gy0 = out + (size_t)batch * modes * samples;
    cx<float>* gy1 = gy0 + (size_t)batch * modes * samples;
    LdbpArgs a = {};
    a.part = reinterpret_cast<double*>(gy1 + (size_t)batch * modes * samples);
    a.in = (const cx<float>*)in;
    a.out = out;
    a.target = (const cx<float>*)target;
    a.fir = (const cx<float>*)params;
    a.kerr = params + 2 * steps * taps;
    a.xs = xs;
    a.ys = ys;
    a.loss = loss;
    a.grad = grad;
    a.J = steps;
    a.K = taps;
    a.KP = kerr_taps;
    a.N = samples;
    a.B = batch;
    double* lslot = a.part + (size_t)steps * (2 * taps + kerr_taps) * batch * (samples / LDBP_CH);
    if (int rc = ldbp_launch_fwd_steps(c, set.fstep, a, modes, lslot)) return rc;
    if (int rc = ldbp_launch_bwd(c, set.bwd, a, modes, gy0, gy1)) return rc;
    return 0;
}

int focss_ldbp_train_step(focss_ctx* c, int steps, int taps, int kerr_taps, int modes, int samples, int batch,
                          float* params, float* adam_m, float* adam_v, int t, float lr, const float* lr_scale,
                          const void* in, const void* target, void* workspace, size_t workspace_bytes,
                          double* loss_host) {
    if (!c || !params || !adam_m || !adam_v || t < 1) return fail(FOCSS_E_ARG, "learned DBP: bad argument");
    if (int rc = ldbp_check(steps, taps, kerr_taps, modes, samples, batch)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    const int np = 2 * steps * taps + steps * kerr_taps;
    // loss and gradient accumulators in the context's buffer (grown, with a
    // synchronisation, when a call needs more)
    const size_t acc_bytes = (size_t)(np + 1) * sizeof(double);
    if (c->ldbp_acc_bytes < acc_bytes) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->ldbp_acc) (void)hipFree(c->ldbp_acc);
        c->ldbp_acc = nullptr;
        c->ldbp_acc_bytes = 0;
        HIP_TRY(dev_malloc(&c->ldbp_acc, acc_bytes));
        c->ldbp_acc_bytes = acc_bytes;
    }
    double* acc = c->ldbp_acc;
    HIP_TRY(hipMemsetAsync(acc, 0, acc_bytes, c->stream));
    int rc = focss_ldbp_gradients(c, steps, taps, kerr_taps, modes, samples, batch, params, in, target, workspace,
                                  workspace_bytes, acc + 1, acc);
    if (rc == 0) {
        hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, c->stream, params, (const double*)(acc + 1),
                           adam_m, adam_v, np, lr, 0.9f, 0.999f, 1e-8f, t, lr_scale);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(FOCSS_E_HIP, std::string("adam: ") + hipGetErrorString(e));
    }
    if (rc == 0 && loss_host) {
        HIP_TRY(hipMemcpyAsync(loss_host, acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return rc;
}

// ---- triplet MLP post-equaliser ----

namespace {

int tmlp_check(int modes, int symbols, int batch) {
    if (modes != 1 && modes != 2) return fail(FOCSS_E_ARG, "triplet MLP: modes must be 1 or 2");
    if (symbols < TM_ROWS || symbols % TM_ROWS != 0 || symbols > (1 << 20))
        return fail(FOCSS_E_ARG, "triplet MLP: symbols must be a multiple of 128 in [128, 2^20]");
    if (batch < 1 || (long long)batch * modes * symbols > (1LL << 27))
        return fail(FOCSS_E_ARG, "triplet MLP: batch must be >= 1 with batch * modes * symbols <= 2^27");
    return 0;
}

int tmlp_splits(int blocks) { return std::min(TM_SPLITS, blocks); }

// workspace: hidden-layer gradient [128][rows] bf16, per-block sums
// [TM_SMALL][blocks] float, weight-gradient partials [splits][128][2240]
// float; each part starts on a 256-byte boundary
size_t tmlp_align(size_t b) { return (b + 255) & ~(size_t)255; }
size_t tmlp_small_offset(size_t rows) { return tmlp_align(rows * TM_H * sizeof(unsigned short)); }
size_t tmlp_part_offset(size_t rows) {
    return tmlp_small_offset(rows) + tmlp_align((size_t)TM_SMALL * (rows / TM_ROWS) * sizeof(float));
}

// the bf16 copy of W1 in the context (fixed size, allocated on first use)
int tmlp_pack(focss_ctx* c, const float* params) {
    if (!c->tmlp_w1b) HIP_TRY(dev_malloc(&c->tmlp_w1b, (size_t)TM_H * TM_KF * sizeof(unsigned short)));
    const int n = TM_H * TM_KF;
    hipLaunchKernelGGL(k_tmlp_pack, dim3((n + 255) / 256), dim3(256), 0, c->stream, params, c->tmlp_w1b, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

TmlpArgs tmlp_args(focss_ctx* c, int modes, int symbols, int batch, const float* params, const void* in) {
    TmlpArgs a = {};
    a.x = (const cx<float>*)in;
    a.w1b = c->tmlp_w1b;
    a.b1 = params + TM_H * TM_KF;
    a.w2 = a.b1 + TM_H;
    a.b2 = a.w2 + 2 * TM_H;
    a.S = symbols;
    a.rows = batch * modes * symbols;
    a.blocks = a.rows / TM_ROWS;
    a.inv_count = 1.0f / (float)a.rows;
    return a;
}

}  // namespace

int focss_tmlp_param_count(void) { return TM_NP; }

size_t focss_tmlp_workspace_bytes(int modes, int symbols, int batch) {
    if (modes < 1 || symbols < TM_ROWS || batch < 1) return 0;
    const size_t rows = (size_t)batch * modes * symbols, blocks = rows / TM_ROWS;
    return tmlp_part_offset(rows) + (size_t)tmlp_splits((int)blocks) * TM_H * TM_KF * sizeof(float);
}

int focss_tmlp_forward(focss_ctx* c, int modes, int symbols, int batch, const float* params, const void* in,
                       void* out, const void* target, double* loss) {
    if (!c || !params || !in || !out || (target && !loss)) return fail(FOCSS_E_ARG, "triplet MLP: null argument");
    if (in == out) return fail(FOCSS_E_ARG, "triplet MLP: in and out must not alias");
    if (int rc = tmlp_check(modes, symbols, batch)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = tmlp_pack(c, params)) return rc;
    TmlpArgs a = tmlp_args(c, modes, symbols, batch, params, in);
    a.out = (cx<float>*)out;
    if (symbols % TI_ROWS == 0) {
        // 256-symbol blocks, two row tiles per wave
        const dim3 grid(a.rows / TI_ROWS);
        if (modes == 1)
            hipLaunchKernelGGL(k_tmlp_infer<1>, grid, dim3(TM_NT), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_tmlp_infer<2>, grid, dim3(TM_NT), 0, c->stream, a);
    } else if (modes == 1) {
        hipLaunchKernelGGL((k_tmlp_fwd<1, false>), dim3(a.blocks), dim3(TM_NT), 0, c->stream, a);
    } else {
        hipLaunchKernelGGL((k_tmlp_fwd<2, false>), dim3(a.blocks), dim3(TM_NT), 0, c->stream, a);
    }
    HIP_TRY(hipGetLastError());
    if (!target) return 0;
    if (!c->tmlp_acc) HIP_TRY(dev_malloc(&c->tmlp_acc, (size_t)(TM_NP + 1) * sizeof(double)));
    HIP_TRY(hipMemsetAsync(c->tmlp_acc, 0, sizeof(double), c->stream));
    const long long n = (long long)a.rows;
    const int grid = (int)std::min<long long>((n + TM_NT - 1) / TM_NT, 4096);
    hipLaunchKernelGGL(k_tmlp_loss, dim3(grid), dim3(TM_NT), 0, c->stream, (const cx<float>*)out,
                       (const cx<float>*)target, n, 1.0 / (double)n, c->tmlp_acc);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(loss, c->tmlp_acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int focss_tmlp_gradients(focss_ctx* c, int modes, int symbols, int batch, const float* params, const void* in,
                         const void* target, void* workspace, size_t workspace_bytes, double* grad, double* loss) {
    if (!c || !params || !in || !target || !workspace || !grad)
        return fail(FOCSS_E_ARG, "triplet MLP: null argument");
    if (int rc = tmlp_check(modes, symbols, batch)) return rc;
    if (workspace_bytes < focss_tmlp_workspace_bytes(modes, symbols, batch))
        return fail(FOCSS_E_ARG, "triplet MLP: workspace too small");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = tmlp_pack(c, params)) return rc;
    TmlpArgs a = tmlp_args(c, modes, symbols, batch, params, in);
    a.target = (const cx<float>*)target;
    a.dpre = (unsigned short*)workspace;
    a.small = (float*)((char*)workspace + tmlp_small_offset((size_t)a.rows));
    a.part = (float*)((char*)workspace + tmlp_part_offset((size_t)a.rows));
    const int splits = tmlp_splits(a.blocks);
    const dim3 g_w1(TM_TP / 32 / 4 + 1, splits);  // 9 x 4 waves cover the 35 triplet blocks
    if (modes == 1) {
        hipLaunchKernelGGL((k_tmlp_fwd<1, true>), dim3(a.blocks), dim3(TM_NT), 0, c->stream, a);
        hipLaunchKernelGGL(k_tmlp_dw1<1>, g_w1, dim3(TM_NT), 0, c->stream, a, splits);
    } else {
        hipLaunchKernelGGL((k_tmlp_fwd<2, true>), dim3(a.blocks), dim3(TM_NT), 0, c->stream, a);
        hipLaunchKernelGGL(k_tmlp_dw1<2>, g_w1, dim3(TM_NT), 0, c->stream, a, splits);
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_tmlp_reduce_w1, dim3((TM_H * TM_KF + 255) / 256), dim3(256), 0, c->stream,
                       (const float*)a.part, splits, grad);
    hipLaunchKernelGGL(k_tmlp_reduce_small, dim3(TM_SMALL), dim3(TM_NT), 0, c->stream, (const float*)a.small,
                       a.blocks, grad, loss);
    HIP_TRY(hipGetLastError());
    return 0;
}

int focss_tmlp_train_step(focss_ctx* c, int modes, int symbols, int batch, float* params, float* adam_m,
                          float* adam_v, int t, float lr, const void* in, const void* target, void* workspace,
                          size_t workspace_bytes, double* loss_host) {
    if (!c || !params || !adam_m || !adam_v || t < 1) return fail(FOCSS_E_ARG, "triplet MLP: bad argument");
    if (int rc = tmlp_check(modes, symbols, batch)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (!c->tmlp_acc) HIP_TRY(dev_malloc(&c->tmlp_acc, (size_t)(TM_NP + 1) * sizeof(double)));
    double* acc = c->tmlp_acc;
    HIP_TRY(hipMemsetAsync(acc, 0, (size_t)(TM_NP + 1) * sizeof(double), c->stream));
    int rc = focss_tmlp_gradients(c, modes, symbols, batch, params, in, target, workspace, workspace_bytes, acc + 1,
                                  acc);
    if (rc == 0) {
        hipLaunchKernelGGL(k_adam, dim3((TM_NP + 255) / 256), dim3(256), 0, c->stream, params,
                           (const double*)(acc + 1), adam_m, adam_v, TM_NP, lr, 0.9f, 0.999f, 1e-8f, t,
                           (const float*)nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(FOCSS_E_HIP, std::string("adam: ") + hipGetErrorString(e));
    }
    if (rc == 0 && loss_host) HIP_TRY(hipMemcpyAsync(loss_host, acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return rc;
}

#define FOCSS_DISPATCH(call_f64, call_f32)                                                   \
    do {                                                                                   \
        if (precision == FOCSS_F64) return call_f64;                                       \
        if (precision == FOCSS_F32) return call_f32;                                       \
        return fail(FOCSS_E_ARG, "precision must be FOCSS_F64 or FOCSS_F32");              \
    } while (0)

int focss_frequency_shift(focss_ctx* c, int precision, int modes, int samples, double dt, double frequency,
                          void* data) {
    if (!c || !data) return fail(FOCSS_E_ARG, "frequency_shift: null argument");
    if (int rc = dsp_check(modes, samples)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((frequency_shift_impl<double>(c, modes, samples, dt, frequency, data)),
                   (frequency_shift_impl<float>(c, modes, samples, dt, frequency, data)));
}

int focss_bandpass_filter(focss_ctx* c, int precision, int modes, int samples, double dt, double frequency,
                          void* data) {
    if (!c || !data) return fail(FOCSS_E_ARG, "bandpass_filter: null argument");
    if (int rc = dsp_check(modes, samples)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((bandpass_impl<double>(c, modes, samples, dt, frequency, data)),
                   (bandpass_impl<float>(c, modes, samples, dt, frequency, data)));
}

int focss_wdm_select(focss_ctx* c, int precision, int modes, int samples, double dt, int channel_index,
                     double channel_bandwidth, const void* in, void* out) {
    if (!c || !in || !out) return fail(FOCSS_E_ARG, "wdm_select: null argument");
    if (int rc = dsp_check(modes, samples)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((wdm_select_impl<double>(c, modes, samples, dt, channel_index, channel_bandwidth, in, out)),
                   (wdm_select_impl<float>(c, modes, samples, dt, channel_index, channel_bandwidth, in, out)));
}

int focss_decimate(focss_ctx* c, int precision, int modes, int samples, int factor, const void* in, void* out) {
    if (!c || !in || !out) return fail(FOCSS_E_ARG, "decimate: null argument");
    if (int rc = dsp_check(modes, samples)) return rc;
    if (factor < 1 || samples % factor || ilog2_int(samples / factor) < 1)
        return fail(FOCSS_E_ARG, "decimate: factor must divide samples, leaving a power of two >= 2");
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((decimate_impl<double>(c, modes, samples, factor, in, out)),
                   (decimate_impl<float>(c, modes, samples, factor, in, out)));
}

int focss_filter(focss_ctx* c, int precision, int modes, int samples, int taps, const double* impulse_response,
                 void* data) {
    if (!c || !data || !impulse_response) return fail(FOCSS_E_ARG, "filter: null argument");
    if (int rc = dsp_check(modes, samples)) return rc;
    if (taps < 1 || taps > samples) return fail(FOCSS_E_ARG, "filter: taps must be in [1, samples]");
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((filter_entry<double>(c, modes, samples, taps, impulse_response, data)),
                   (filter_entry<float>(c, modes, samples, taps, impulse_response, data)));
}

// ---- constant equalisers (src/focss/equalizer/linear.h) ----

namespace {

// st = {sum |d|^2, sum |a|^2, Re sum conj(a) d, Im sum conj(a) d} (host), synchronous
int linear_stats(focss_ctx* c, int precision, int modes, int samples, const void* desired, const void* actual,
                 double st[4]) {
    if (!c || !desired || !actual) return fail(FOCSS_E_ARG, "linear equaliser: null argument");
    if (modes < 1 || samples < 1 || (long long)modes * samples > (1LL << 31))
        return fail(FOCSS_E_ARG, "linear equaliser: modes and samples must be >= 1");
    if (precision != FOCSS_F64 && precision != FOCSS_F32)
        return fail(FOCSS_E_ARG, "precision must be FOCSS_F64 or FOCSS_F32");
    HIP_TRY(hipSetDevice(c->device));
    if (!c->lin_stats) HIP_TRY(dev_malloc(&c->lin_stats, 4 * sizeof(double)));
    HIP_TRY(hipMemsetAsync(c->lin_stats, 0, 4 * sizeof(double), c->stream));
    const long long n = (long long)modes * samples;
    const int grid = (int)std::min<long long>((n + LS_NT - 1) / LS_NT, 1024);
    if (precision == FOCSS_F64)
        hipLaunchKernelGGL(k_linear_stats<double>, dim3(grid), dim3(LS_NT), 0, c->stream,
                           (const cx<double>*)desired, (const cx<double>*)actual, n, c->lin_stats);
    else
        hipLaunchKernelGGL(k_linear_stats<float>, dim3(grid), dim3(LS_NT), 0, c->stream,
                           (const cx<float>*)desired, (const cx<float>*)actual, n, c->lin_stats);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(st, c->lin_stats, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

// phase_shift_equalizer.cpp:30-46: q2 of actual e^{i angle} against desired
// for angle = 2 pi i / angle_steps, the first maximum (starting from the
// unrotated field).  sum |d - a e^{i t}|^2 = sum |d|^2 + sum |a|^2
// - 2 Re(e^{-i t} sum d conj(a)), so one pass over the fields gives every
// candidate's q2.
int focss_pse_estimate(focss_ctx* c, int precision, int modes, int samples, const void* desired, const void* actual,
                       int angle_steps, double* angle) {
    if (!angle) return fail(FOCSS_E_ARG, "pse_estimate: null argument");
    if (angle_steps < 0) return fail(FOCSS_E_ARG, "pse_estimate: angle_steps must be >= 0");
    double st[4];
    if (int rc = linear_stats(c, precision, modes, samples, desired, actual, st)) return rc;
    const double pd = st[0], pa = st[1], cr = st[2], ci = st[3];  // C = sum d conj(a) = sum conj(a) d
    auto q2 = [&](double t) {
        const double num = pd + pa - 2.0 * (cr * std::cos(t) + ci * std::sin(t));
        return num > 0.0 ? -10.0 * std::log10(num / pd) : HUGE_VAL;
    };
    const double two_pi = 6.283185307179586476925286766559;
    double best = q2(0.0), arg = 0.0;
    for (int i = 0; i < angle_steps; ++i) {
        const double t = two_pi * (double)i / angle_steps;
        const double q = q2(t);
        if (best < q) {
            best = q;
            arg = t;
        }
    }
    *angle = arg;
    return 0;
}

// scalar_equalizer.cpp:22-33: sum conj(a) d / sum |a|^2 (complex, [re, im])
int focss_scalar_estimate(focss_ctx* c, int precision, int modes, int samples, const void* desired,
                          const void* actual, double* scalar) {
    if (!scalar) return fail(FOCSS_E_ARG, "scalar_estimate: null argument");
    double st[4];
    if (int rc = linear_stats(c, precision, modes, samples, desired, actual, st)) return rc;
    scalar[0] = st[2] / st[1];
    scalar[1] = st[3] / st[1];
    return 0;
}

// homothety_equalizer.cpp:22-33: |sum conj(a) d / sum |a|^2|
int focss_homothety_estimate(focss_ctx* c, int precision, int modes, int samples, const void* desired,
                             const void* actual, double* scalar) {
    if (!scalar) return fail(FOCSS_E_ARG, "homothety_estimate: null argument");
    double st[4];
    if (int rc = linear_stats(c, precision, modes, samples, desired, actual, st)) return rc;
    *scalar = std::hypot(st[2] / st[1], st[3] / st[1]);
    return 0;
}

int focss_rrc_shaping(focss_ctx* c, int precision, int modes, int n_symbols, const void* symbols, int osf,
                      double roll_off, double gain, void* out) {
    if (!c || !symbols || !out) return fail(FOCSS_E_ARG, "rrc_shaping: null argument");
    if (osf < 1 || !(roll_off > 0 && roll_off <= 1)) return fail(FOCSS_E_ARG, "rrc_shaping: bad osf or roll-off");
    if (int rc = dsp_check(modes, n_symbols * osf)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((rrc_shaping_impl<double>(c, modes, n_symbols, symbols, osf, roll_off, gain, out)),
                   (rrc_shaping_impl<float>(c, modes, n_symbols, symbols, osf, roll_off, gain, out)));
}

int focss_rrc_sampling(focss_ctx* c, int precision, int modes, int samples, const void* signal, int osf,
                       double roll_off, double gain, void* out) {
    if (!c || !signal || !out) return fail(FOCSS_E_ARG, "rrc_sampling: null argument");
    if (osf < 1 || samples % osf || !(roll_off > 0 && roll_off <= 1))
        return fail(FOCSS_E_ARG, "rrc_sampling: bad osf or roll-off");
    if (int rc = dsp_check(modes, samples)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((rrc_sampling_impl<double>(c, modes, samples, signal, osf, roll_off, gain, out)),
                   (rrc_sampling_impl<float>(c, modes, samples, signal, osf, roll_off, gain, out)));
}

int focss_field_from_records(focss_ctx* c, int precision, int envelopes, int samples, const void* records,
                             void* field) {
    if (int rc = records_check(c, envelopes, samples, records, field, "field_from_records")) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((envelope_transpose<double>(c, true, envelopes, samples, records, field)),
                   (envelope_transpose<float>(c, true, envelopes, samples, records, field)));
}

int focss_field_to_records(focss_ctx* c, int precision, int envelopes, int samples, const void* field,
                           void* records) {
    if (int rc = records_check(c, envelopes, samples, field, records, "field_to_records")) return rc;
    HIP_TRY(hipSetDevice(c->device));
    FOCSS_DISPATCH((envelope_transpose<double>(c, false, envelopes, samples, field, records)),
                   (envelope_transpose<float>(c, false, envelopes, samples, field, records)));
}

#undef FOCSS_DISPATCH

int focss_set_tile(focss_ctx* c, int precision, int log2_samples, int variant) {
    if (!c || precision < 0 || precision > 1 || log2_samples < 0 || log2_samples > 20 || variant < 0 ||
        variant >= kTileVariants)
        return fail(FOCSS_E_ARG, "bad tile selection");
    c->tile_variant[0][precision][log2_samples] = variant;
    c->tile_variant[1][precision][log2_samples] = variant;
    drop_graph(c);
    return 0;
}

int focss_set_tile_pass(focss_ctx* c, int precision, int log2_samples, int pass, int variant) {
    if (!c || precision < 0 || precision > 1 || log2_samples < 0 || log2_samples > 20 || pass < 0 || pass > 1 ||
        variant < 0 || variant >= kTileVariants)
        return fail(FOCSS_E_ARG, "bad tile selection");
    c->tile_variant[pass][precision][log2_samples] = variant;
    drop_graph(c);
    return 0;
}

int focss_tile_variants(int precision, int log2_samples) {
    // variants are numbered from 0; a pass without variant v runs its first
    // usable one
    int n = 0;
    for (int v = 0; v < kTileVariants; ++v)
        if (time_ok(precision, log2_samples, 1, v) || spec_ok(precision, log2_samples, v)) n = v + 1;
    return n;
}

}  // extern "C"
