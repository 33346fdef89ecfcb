// Vector-ALU issue rates of this GPU, for the rooflines of the force sweeps.
//
// Each rate kernel runs independent dependency chains (8 per work-item, 16 work-groups of
// 256 per CU) of one instruction kind with no memory traffic, so its duration is set by
// instruction issue alone; red_valu_rates reports lane-instructions per second.  The pair
// kernel repeats the mixed-precision sweep's pair sequence (FP64 differences, conversions,
// FP32 r^2, v_rsq_f32, m r^-3, FP32 accumulation and the screen's minimum; sources by scalar
// loads from a 4096-source table that stays in cache) and reports pairs per second: the ALU
// ceiling of that sequence on this GPU; the FP64 pair kernel does the same for the FP64
// sweep's pair term (IEEE square root and quotient, sums in ascending j).
#include "fp64_core.hpp"
#include "red_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using red::d4;

namespace {

constexpr int RCH = 8;            // chains per work-item
constexpr int RATE_BLOCKS = 256 * 16;
constexpr int RATE_THREADS = 256;
constexpr int PAIR_SRC = 4096;

enum RateKind { R_FMA32, R_MUL32, R_ADD64, R_CVT, R_RSQ32, R_SPLIT, R_KINDS };

template <int KIND>
__global__ __launch_bounds__(256) void k_rate(float* __restrict__ out, int iters, float b, float c)
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float s = 0.0f;
    // FP64 chains for the FP64 kinds and for the odd wavefronts of the split kind
    const bool f64 = KIND == R_ADD64 || KIND == R_CVT || (KIND == R_SPLIT && (w & 1));
    if (f64) {
        double d[RCH];
        const double e = (double)b * 1e-9;
        for (int k = 0; k < RCH; k++) d[k] = threadIdx.x * 1e-3 + k + b;
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int k = 0; k < RCH; k++) d[k] = KIND == R_CVT ? (double)(float)d[k] : d[k] + e;
        }
        for (int k = 0; k < RCH; k++) s += (float)d[k];
    } else {
        float a[RCH];
        for (int k = 0; k < RCH; k++) a[k] = 1.0f + threadIdx.x * 1e-3f + k + b;
        for (int it = 0; it < iters; it++) {
#pragma unroll
            for (int k = 0; k < RCH; k++) {
                if (KIND == R_MUL32) a[k] = a[k] * b;
                else if (KIND == R_RSQ32) a[k] = __builtin_amdgcn_rsqf(a[k]);
                else a[k] = __builtin_fmaf(a[k], b, c);
            }
        }
        for (int k = 0; k < RCH; k++) s += a[k];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// The mixed sweep's pair sequence (k_grav_mixed_bcast, two sinks per work-item, collision
// screen on) over `reps` passes of the source table.
__global__ __launch_bounds__(256) void k_rate_pair(const d4* __restrict__ src, float* __restrict__ out, int reps)
{
    constexpr int SPT = 2;
    double xi[SPT], yi[SPT], zi[SPT];
    float sx[SPT], sy[SPT], sz[SPT], mn[SPT];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        xi[s] = 0.37 + 1e-6 * (t * SPT + s);
        yi[s] = -1.21 + 3e-7 * t + 1e-8 * s;
        zi[s] = 0.01 * s + 1e-9 * t;
        sx[s] = sy[s] = sz[s] = 0.0f;
        mn[s] = 3.0e38f;
    }
    for (int rep = 0; rep < reps; rep++) {
#pragma unroll 4
        for (int j = 0; j < PAIR_SRC; ++j) {
            const d4 q = src[j];
            const float qm = __uint_as_float((unsigned)((unsigned long long)__double_as_longlong(q.w) & 0xffffffffull));
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                const float dx = (float)(q.x - xi[s]);
                const float dy = (float)(q.y - yi[s]);
                const float dz = (float)(q.z - zi[s]);
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float sc = (qm * inv) * (inv * inv);
                sx[s] = __builtin_fmaf(sc, dx, sx[s]);
                sy[s] = __builtin_fmaf(sc, dy, sy[s]);
                sz[s] = __builtin_fmaf(sc, dz, sz[s]);
                mn[s] = fminf(mn[s], r2);
            }
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int s = 0; s < SPT; s++) acc += (sx[s] + sy[s]) + (sz[s] + mn[s]);
    out[t] = acc;
}

// The FP64 sweep's pair term in the reference's operation order (differences, r^2, the IEEE
// square root and quotient through their core sequences, the three products), added in
// ascending j into the work-item's own sums: two sinks per work-item, sources by scalar loads.
__global__ __launch_bounds__(256) void k_rate_fp64_pair(const d4* __restrict__ src, float* __restrict__ out, int reps)
{
    constexpr int SPT = 2;
    double xi[SPT], yi[SPT], zi[SPT], ax[SPT], ay[SPT], az[SPT];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int s = 0; s < SPT; s++) {
        xi[s] = 0.37 + 1e-6 * (t * SPT + s);
        yi[s] = -1.21 + 3e-7 * t + 1e-8 * s;
        zi[s] = 0.01 * s + 1e-9 * t;
        ax[s] = ay[s] = az[s] = 0.0;
    }
    for (int rep = 0; rep < reps; rep++) {
#pragma unroll 4
        for (int j = 0; j < PAIR_SRC; ++j) {
            const d4 q = src[j];
#pragma unroll
            for (int s = 0; s < SPT; s++) {
                const double dx = q.x - xi[s];
                const double dy = q.y - yi[s];
                const double dz = q.z - zi[s];
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const double d = red::sqrt_core(r2);
                const double sc = red::div_core(q.w, d * r2);
                ax[s] += sc * dx;
                ay[s] += sc * dy;
                az[s] += sc * dz;
            }
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < SPT; s++) acc += (ax[s] + ay[s]) + az[s];
    out[t] = (float)acc;
}

// Times launch() `reps` times after one warm-up launch; returns the fastest, in seconds.
template <class F>
hipError_t time_best(F launch, int reps, double& best)
{
    hipEvent_t a = nullptr, b = nullptr;
    RED_HIP_TRY(hipEventCreate(&a));
    RED_HIP_TRY(hipEventCreate(&b));
    hipError_t e = hipSuccess;
    best = 1e300;
    launch();
    for (int k = 0; k < reps && e == hipSuccess; k++) {
        e = hipEventRecord(a, 0);
        if (e == hipSuccess) launch();
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(b, 0);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess) best = std::min(best, 1e-3 * ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e;
}

}  // namespace

extern "C" {

int red_valu_rates(int device, double* out, int n_out)
{
    if (!out || n_out < RED_VALU_RATES) return RED_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return RED_ERR_HIP;
    float* buf = nullptr;
    d4* src = nullptr;
    if (hipMalloc((void**)&buf, sizeof(float) * RATE_BLOCKS * RATE_THREADS) != hipSuccess) return RED_ERR_HIP;
    if (hipMalloc((void**)&src, sizeof(d4) * PAIR_SRC) != hipSuccess) {
        (void)hipFree(buf);
        return RED_ERR_HIP;
    }
    // sources spread over a disk of 5 AU with FP32 masses of 1e-10 in the low word of w
    std::vector<d4> h(PAIR_SRC);
    for (int j = 0; j < PAIR_SRC; j++) {
        const double ang = 0.001534 * j, rad = 0.4 + 4.6 * ((j * 2654435761u) % 4096) / 4096.0;
        const float m = 1e-10f;
        unsigned mb;
        __builtin_memcpy(&mb, &m, sizeof(mb));
        const unsigned long long wb = mb;
        double w;
        __builtin_memcpy(&w, &wb, sizeof(w));
        h[j] = d4{rad * std::cos(ang), rad * std::sin(ang), 0.001 * ((j % 7) - 3), w};
    }
    hipError_t e = hipMemcpy(src, h.data(), sizeof(d4) * PAIR_SRC, hipMemcpyHostToDevice);
    const int iters = 16384;
    const double lanes = (double)RATE_BLOCKS * RATE_THREADS;
    const dim3 g(RATE_BLOCKS), b(RATE_THREADS);
    double s = 0.0;
    // about 50 ms of ALU work first, so that the clocks have settled before anything is timed
    for (int k = 0; k < 16 && e == hipSuccess; k++) {
        hipLaunchKernelGGL(k_rate_pair, g, b, 0, 0, src, buf, 2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
#define RATE(KIND, PER)                                                                                     \
    if (e == hipSuccess) {                                                                                  \
        e = time_best([&] { hipLaunchKernelGGL(k_rate<KIND>, g, b, 0, 0, buf, iters, 0.999f, 1e-3f); }, 5, s); \
        out[KIND] = lanes * iters * RCH * (PER) / s;                                                        \
    }
    RATE(R_FMA32, 1)
    RATE(R_MUL32, 1)
    RATE(R_ADD64, 1)
    RATE(R_CVT, 2)
    RATE(R_RSQ32, 1)
    RATE(R_SPLIT, 1)
#undef RATE
    if (e == hipSuccess) {
        const int reps = 2;
        e = time_best([&] { hipLaunchKernelGGL(k_rate_pair, g, b, 0, 0, src, buf, reps); }, 5, s);
        out[R_KINDS] = lanes * 2.0 * PAIR_SRC * reps / s;
    }
    if (e == hipSuccess) {
        // FP64 masses for the FP64 terms (the table's w holds an FP32 mass in its low word)
        for (int j = 0; j < PAIR_SRC; j++) h[j].w = 1e-10 * (1.0 + (j % 13) * 0.01);
        e = hipMemcpy(src, h.data(), sizeof(d4) * PAIR_SRC, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        const int reps = 1;
        e = time_best([&] { hipLaunchKernelGGL(k_rate_fp64_pair, g, b, 0, 0, src, buf, reps); }, 5, s);
        out[R_KINDS + 1] = lanes * 2.0 * PAIR_SRC * reps / s;
    }
    (void)hipFree(src);
    (void)hipFree(buf);
    return e == hipSuccess ? RED_OK : RED_ERR_HIP;
}

}  // extern "C"
