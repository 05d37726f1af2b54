// FP64 issue-cost probe: per-lane throughput of FMA, IEEE division, the hardware reciprocal and
// reciprocal square root, and square root, each relative to FMA. Every thread runs 8 independent
// chains of 256 operations, so latency is hidden and the times measure issue cost. Used to
// price the sweeps' divisions.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

namespace {

constexpr int kIters = 256;
constexpr int kChains = 8;

enum Op { OP_FMA, OP_DIV, OP_RCP, OP_SQRT, OP_RSQ, OP_MUL_ADD, OP_COUNT };
const char* const kNames[OP_COUNT] = {"fma", "ieee div", "add + rcp", "add + sqrt", "add + rsq", "mul + add"};

template <int OP>
__global__ void __launch_bounds__(256) k_rate(const double* __restrict__ a, const double* __restrict__ b,
                                              double* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[kChains], y[kChains];
#pragma unroll
    for (int j = 0; j < kChains; j++) {
        x[j] = a[i] + j;
        y[j] = b[i] + 0.5 * j;
    }
    for (int it = 0; it < kIters; it++) {
#pragma unroll
        for (int j = 0; j < kChains; j++) {
            if constexpr (OP == OP_FMA) x[j] = __builtin_fma(x[j], y[j], 0.25);
            else if constexpr (OP == OP_DIV) x[j] = y[j] / x[j];
            else if constexpr (OP == OP_RCP) x[j] = __builtin_amdgcn_rcp(x[j] + y[j]);
            else if constexpr (OP == OP_SQRT) x[j] = __builtin_sqrt(x[j] + y[j]);
            else if constexpr (OP == OP_RSQ) x[j] = __builtin_amdgcn_rsq(x[j] + y[j]);
            else x[j] = x[j] * y[j] + 1.0;
        }
    }
    double s = 0;
#pragma unroll
    for (int j = 0; j < kChains; j++) s += x[j];
    out[i] = s;
}

template <int OP>
float time_ms(const double* a, const double* b, double* o, int n, int reps)
{
    hipEvent_t e0, e1;
    (void) hipEventCreate(&e0);
    (void) hipEventCreate(&e1);
    const dim3 grid((n + 255) / 256);
    k_rate<OP><<<grid, 256>>>(a, b, o, n);
    (void) hipEventRecord(e0);
    for (int r = 0; r < reps; r++) k_rate<OP><<<grid, 256>>>(a, b, o, n);
    (void) hipEventRecord(e1);
    (void) hipEventSynchronize(e1);
    float ms = 0;
    (void) hipEventElapsedTime(&ms, e0, e1);
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    return ms / reps;
}

}  // namespace

int main()
{
    const int n = 256 * 4 * 16 * 64;   // 16 waves for each SIMD of 256 CUs
    std::vector<double> ha(n), hb(n);
    for (int i = 0; i < n; i++) {
        ha[i] = 1.0 + (i % 97) * 0.01;
        hb[i] = 2.0 + (i % 89) * 0.01;
    }
    double *a = nullptr, *b = nullptr, *o = nullptr;
    CHECK(hipMalloc((void**) &a, n * sizeof(double)));
    CHECK(hipMalloc((void**) &b, n * sizeof(double)));
    CHECK(hipMalloc((void**) &o, n * sizeof(double)));
    CHECK(hipMemcpy(a, ha.data(), n * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(b, hb.data(), n * sizeof(double), hipMemcpyHostToDevice));
    float t[OP_COUNT];
    t[OP_FMA] = time_ms<OP_FMA>(a, b, o, n, 20);
    t[OP_DIV] = time_ms<OP_DIV>(a, b, o, n, 20);
    t[OP_RCP] = time_ms<OP_RCP>(a, b, o, n, 20);
    t[OP_SQRT] = time_ms<OP_SQRT>(a, b, o, n, 20);
    t[OP_RSQ] = time_ms<OP_RSQ>(a, b, o, n, 20);
    t[OP_MUL_ADD] = time_ms<OP_MUL_ADD>(a, b, o, n, 20);
    CHECK(hipDeviceSynchronize());
    const double ops = (double) n * kIters * kChains;
    const double slot = t[OP_MUL_ADD] / 2;   // one full-rate FP64 instruction
    std::printf("%-11s %9s %16s %22s\n", "operation", "ms", "G lane-ops/s", "full-rate instr slots");
    for (int k = 0; k < OP_COUNT; k++)
        std::printf("%-11s %9.3f %16.1f %22.2f\n", kNames[k], t[k], ops / (t[k] * 1e-3) / 1e9, t[k] / slot);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(o));
    return 0;
}
