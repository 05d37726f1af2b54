// Access-pattern probe of the Y march: a copy with no arithmetic that walks memory as the march
// does (4 state arrays read, 4 written, (ny + 8) x (nx + 8) doubles at 8192^2; each thread one
// column; a block of BT threads covers BT adjacent columns over a run of 128 rows plus 4 warm-up
// rows each side; runs numbered row band by row band, spread over the XCDs), with the block's
// waves either marching independently or meeting a block barrier every SYNC rows. Prints the
// time and the algorithmic rate (4 words in + 4 out per real cell).
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

namespace {

constexpr int G = 4, NX = 8192, NY = 8192, RL = NX + 2 * G, L = 128;

struct Arrays {
    const double* in[4];
    double* out[4];
};

// blocks of one XCD take a contiguous share of the run list (blocks are dealt round-robin over 8)
__device__ __forceinline__ int xcd_block(int b, int n)
{
    const int per = n / 8;
    return b < per * 8 ? (b % 8) * per + b / 8 : b;
}

template <int BT, int SYNC>
__global__ void __launch_bounds__(BT) k_ymarch_copy(Arrays a, int ncolblk)
{
    const int bid = xcd_block(blockIdx.x, gridDim.x);
    const int seg = bid / ncolblk;
    const int y0 = seg * L, y1 = min(y0 + L, NY);
    const int x = (bid - seg * ncolblk) * BT + (int) threadIdx.x;
    const bool ok = x < NX;
    double acc[4] = {0, 0, 0, 0};
    int k = 0;
    for (int n = y0 - G; n < y1 + G; n++) {   // rows n + G in [y0, y1 + 2G) of the NY + 2G
        const long idx = (long) (n + G) * RL + G + (ok ? x : 0);
        if (ok)
            for (int q = 0; q < 4; q++) acc[q] += a.in[q][idx];
        if (ok && n >= y0 && n < y1)
            for (int q = 0; q < 4; q++) a.out[q][idx] = acc[q];
        if constexpr (SYNC > 0) {
            if (++k == SYNC) {
                k = 0;
                __syncthreads();
            }
        }
    }
}

template <int BT, int SYNC>
int run(const Arrays& a, const char* name)
{
    const int ncolblk = (NX + BT - 1) / BT, nseg = (NY + L - 1) / L;
    const dim3 grid((unsigned) (ncolblk * nseg));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    k_ymarch_copy<BT, SYNC><<<grid, BT>>>(a, ncolblk);
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < 10; r++) k_ymarch_copy<BT, SYNC><<<grid, BT>>>(a, ncolblk);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= 10;
    std::printf("%-34s %8.3f ms  %7.0f GB/s\n", name, ms, 64.0 * NX * NY / (ms * 1e-3) / 1e9);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return 0;
}

}  // namespace

int main()
{
    const size_t n = (size_t) (NY + 2 * G) * RL;
    Arrays a;
    double* bufs[8];
    for (int q = 0; q < 8; q++) {
        CHECK(hipMalloc((void**) &bufs[q], n * sizeof(double)));
        CHECK(hipMemset(bufs[q], 0, n * sizeof(double)));
    }
    for (int q = 0; q < 4; q++) {
        a.in[q] = bufs[q];
        a.out[q] = bufs[4 + q];
    }
    std::printf("%-34s %8s     %s\n", "block, synchronisation", "time", "algorithmic rate");
    int rc = run<256, 0>(a, "256 threads, independent waves");
    rc |= run<256, 1>(a, "256 threads, barrier every row");
    rc |= run<256, 2>(a, "256 threads, barrier every 2 rows");
    rc |= run<256, 4>(a, "256 threads, barrier every 4 rows");
    rc |= run<256, 8>(a, "256 threads, barrier every 8 rows");
    rc |= run<512, 0>(a, "512 threads, independent waves");
    rc |= run<512, 1>(a, "512 threads, barrier every row");
    rc |= run<1024, 1>(a, "1024 threads, barrier every row");
    CHECK(hipDeviceSynchronize());
    for (int q = 0; q < 8; q++) CHECK(hipFree(bufs[q]));
    return rc;
}
