// Self-tests of the shortened FP64 sequences of device_math.hpp, on the device: the
// shared-divisor quotients (with_divisor: three quotients over one divisor against the
// compiler's IEEE a / b) and the square root (sqrt_t<double> against the compiler's
// __builtin_sqrt), bit for bit, for operands drawn over the whole FP64 range (subnormals, zeros
// of both signs, infinities, NaNs), around the bounds of the ranges where the short sequences
// are taken, and in the range of physical states. Built as its own library,
// libarmon_selftest.so; the solver library does not contain it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_math.hpp"

namespace {

__device__ uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
__device__ uint64_t to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// A double with unbiased exponent e (normal range) or biased exponent 0 (subnormal) when e is
// below -1022, a random mantissa (or zero, one time in eight) and the given sign.
__device__ double with_exponent(uint64_t& s, int e, bool negative)
{
    const uint64_t r = splitmix64(s);
    const uint64_t mant = (r & 7) == 0 ? 0 : (r >> 12);
    const uint64_t biased = e < -1022 ? 0 : (uint64_t) (e + 1023);
    return from_bits(((uint64_t) negative << 63) | (biased << 52) | (mant & 0xFFFFFFFFFFFFFull));
}

__device__ double special(uint64_t r)
{
    switch (r % 8) {
    case 0: return 0.0;
    case 1: return -0.0;
    case 2: return __builtin_huge_val();
    case 3: return -__builtin_huge_val();
    case 4: return __builtin_nan("");
    case 5: return from_bits(1);                       // smallest subnormal
    case 6: return from_bits(0x000FFFFFFFFFFFFFull);   // largest subnormal
    default: return from_bits(0x7FEFFFFFFFFFFFFFull);  // largest finite
    }
}

__device__ double draw_divisor(uint64_t& s)
{
    const uint64_t r = splitmix64(s);
    const bool neg = (r >> 8) & 1;
    const int span = (int) ((r >> 16) & 0xFFFF);
    switch (r & 7) {
    case 0: return with_exponent(s, -1075 + span % 2099, neg);   // every exponent, both signs
    case 1: return with_exponent(s, -210 + span % 421, false);   // around the divisor range
    case 2: {                                                     // at its bounds
        const int e[6] = {-201, -200, -199, 199, 200, 201};
        return with_exponent(s, e[span % 6], false);
    }
    case 3: return special(r >> 20);
    default: return with_exponent(s, -20 + span % 41, false);   // physical states
    }
}

__device__ double draw_numerator(uint64_t& s)
{
    const uint64_t r = splitmix64(s);
    const bool neg = (r >> 8) & 1;
    const int span = (int) ((r >> 16) & 0xFFFF);
    switch (r & 7) {
    case 0: return with_exponent(s, -1075 + span % 2099, neg);
    case 1: return with_exponent(s, -610 + span % 921, neg);     // around the numerator range
    case 2: {
        const int e[6] = {-601, -600, -599, 299, 300, 301};
        return with_exponent(s, e[span % 6], neg);
    }
    case 3: return neg ? -0.0 : 0.0;
    case 4: return special(r >> 20);
    default: return with_exponent(s, -30 + span % 61, neg);
    }
}

// counts[0]: quotients whose bits differ from a / b; counts[1]: triples that took the shared
// sequence (every operand inside its range); first[0..3]: b, a0, a1, a2 of a differing triple.
__global__ void __launch_bounds__(256) k_check_shared_divisor(uint64_t seed, int64_t n, unsigned long long* counts,
                                                               double* first, int* first_taken)
{
    unsigned long long bad = 0, fast = 0;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        uint64_t s = seed ^ ((uint64_t) i * 0xD1B54A32D192ED03ull);
        const double b = draw_divisor(s);
        const double a0 = draw_numerator(s), a1 = draw_numerator(s), a2 = draw_numerator(s);
        double q0, q1, q2;
        armon::with_divisor(b, [&](auto div) {
            q0 = div(a0);
            q1 = div(a1);
            q2 = div(a2);
        });
        const int miss = (int) (to_bits(q0) != to_bits(a0 / b)) + (int) (to_bits(q1) != to_bits(a1 / b))
                + (int) (to_bits(q2) != to_bits(a2 / b));
        bad += (unsigned long long) miss;
        fast += (unsigned long long) (armon::shared_divisor_ok(b) & armon::shared_numerator_ok(a0)
                                      & armon::shared_numerator_ok(a1) & armon::shared_numerator_ok(a2));
        if (miss && atomicCAS(first_taken, 0, 1) == 0) {
            first[0] = b;
            first[1] = a0;
            first[2] = a1;
            first[3] = a2;
        }
    }
    if (bad) atomicAdd(&counts[0], bad);
    if (fast) atomicAdd(&counts[1], fast);
}

__device__ double draw_radicand(uint64_t& s)
{
    const uint64_t r = splitmix64(s);
    const bool neg = ((r >> 8) & 15) == 0;
    const int span = (int) ((r >> 16) & 0xFFFF);
    switch (r & 7) {
    case 0: return with_exponent(s, -1075 + span % 2099, neg);   // every exponent
    case 1: return with_exponent(s, -777 + span % 21, false);    // around 2^-767
    case 2: return special(r >> 20);
    case 3: return with_exponent(s, 1013 + span % 11, false);    // near the largest finite
    default: return with_exponent(s, -30 + span % 61, false);    // physical states
    }
}

// counts[0]: roots whose bits differ from __builtin_sqrt; counts[1]: operands that took the
// short sequence; first[0]: the operand of a differing root.
__global__ void __launch_bounds__(256) k_check_sqrt(uint64_t seed, int64_t n, unsigned long long* counts,
                                                     double* first, int* first_taken)
{
    unsigned long long bad = 0, fast = 0;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        uint64_t s = seed ^ ((uint64_t) i * 0xD1B54A32D192ED03ull);
        const double x = draw_radicand(s);
        const int miss = (int) (to_bits(armon::sqrt_t<double>(x)) != to_bits(__builtin_sqrt(x)));
        bad += (unsigned long long) miss;
        fast += (unsigned long long) armon::sqrt_core_ok(x);
        if (miss && atomicCAS(first_taken, 0, 1) == 0) first[0] = x;
    }
    if (bad) atomicAdd(&counts[0], bad);
    if (fast) atomicAdd(&counts[1], fast);
}

using CheckKernel = void (*)(uint64_t, int64_t, unsigned long long*, double*, int*);

// Runs one check kernel over n samples and copies its two counts and, when a sample failed,
// its nfirst operands back. Returns 0, or the HIP error code.
int run_check(CheckKernel kernel, uint64_t seed, int64_t n, uint64_t* mismatches, uint64_t* taken_short,
              double* first_bad, int nfirst)
{
    unsigned long long* counts = nullptr;
    double* first = nullptr;
    int* taken = nullptr;
    hipError_t e = hipMalloc((void**) &counts, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**) &first, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**) &taken, sizeof(int));
    if (e == hipSuccess) e = hipMemset(counts, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(taken, 0, sizeof(int));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kernel, dim3(2048), dim3(256), 0, 0, seed, n, counts, first, taken);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    unsigned long long h[2] = {0, 0};
    int t = 0;
    if (e == hipSuccess) e = hipMemcpy(h, counts, sizeof(h), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&t, taken, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && t) e = hipMemcpy(first_bad, first, nfirst * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        *mismatches = h[0];
        *taken_short = h[1];
    }
    (void) hipFree(counts);
    (void) hipFree(first);
    (void) hipFree(taken);
    return (int) e;
}

}  // namespace

// Checks n triples drawn from seed. mismatches: quotients that differ from a / b; shared:
// triples that took the shared sequence; first_bad: b, a0, a1, a2 of one differing triple
// (untouched when none differs). Returns 0, or the HIP error code.
extern "C" __attribute__((visibility("default"))) int armon_selftest_shared_divisor(
        uint64_t seed, int64_t n, uint64_t* mismatches, uint64_t* shared, double* first_bad)
{
    return run_check(k_check_shared_divisor, seed, n, mismatches, shared, first_bad, 4);
}

// Checks n square roots drawn from seed. mismatches: roots that differ from __builtin_sqrt;
// fast: operands that took the short sequence; first_bad: the operand of one differing root
// (untouched when none differs). Returns 0, or the HIP error code.
extern "C" __attribute__((visibility("default"))) int armon_selftest_sqrt(uint64_t seed, int64_t n,
                                                                          uint64_t* mismatches, uint64_t* fast,
                                                                          double* first_bad)
{
    return run_check(k_check_sqrt, seed, n, mismatches, fast, first_bad, 1);
}
