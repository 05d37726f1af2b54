// The 14 drop-in entry points of the Armon kernel C-ABI (reference src/kernels/kernels.h:10-86),
// each one HIP kernel over its Range, plus the ABI helpers and the library plumbing.
//
// Launch geometry: a Range scaled through an InnerRange2D is a set of row segments, so the
// kernels use a 2D grid (x = column inside the row segment, y = row) instead of the
// reference's per-element 64-bit div/mod (indexing.h:67-72). Consecutive lanes touch
// consecutive addresses.
#include "internal.hpp"
#include "device_math.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace armon;

// ---------------------------------------------------------------------------------------
// Error funnel (utils.cpp:8-17, 49-54)
// ---------------------------------------------------------------------------------------

static void default_exception_handler(const char* kernel, const char* msg)
{
    printf("Exception in kernel %s: %s\n", kernel, msg);
    exit(1);
}

extern "C" ARMON_EXPORT void (*raise_exception_handler)(const char* kernel, const char* msg) = default_exception_handler;

static thread_local std::string g_last_error;

namespace armon {

static uint64_t g_report_count = 0;

uint64_t report_count() { return g_report_count; }

void report(const char* kernel, const Error& e)
{
    g_report_count++;
    std::string msg = "(" + e.type + ") " + e.what();
    g_last_error = std::string(kernel) + ": " + msg;
    raise_exception_handler(kernel, msg.c_str());
}

void report_std(const char* kernel, const std::exception& e)
{
    report(kernel, Error("std::exception", e.what()));
}

// ---------------------------------------------------------------------------------------
// Stream and scratch
// ---------------------------------------------------------------------------------------

static hipStream_t g_stream = nullptr;

hipStream_t stream() { return g_stream; }

Scratch& scratch()
{
    static std::mutex mtx;
    static std::vector<Scratch> per_device;
    std::lock_guard<std::mutex> lk(mtx);
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if ((int) per_device.size() <= dev) per_device.resize(dev + 1);
    Scratch& s = per_device[dev];
    if (!s.partials) {
        void* p = nullptr;
        size_t bytes = sizeof(double) * Scratch::kPartials + 256 + sizeof(double) * 8;
        hip_check(hipMalloc(&p, bytes), "hipMalloc(scratch)");
        char* base = static_cast<char*>(p);
        s.partials = reinterpret_cast<double*>(base);
        s.min_bits = reinterpret_cast<unsigned long long*>(base + sizeof(double) * Scratch::kPartials);
        s.result = reinterpret_cast<double*>(base + sizeof(double) * Scratch::kPartials + 256);
    }
    return s;
}

// ---------------------------------------------------------------------------------------
// Index spans and checks
// ---------------------------------------------------------------------------------------

Iter2D make_iter2d(const Range& r, const InnerRange2D& ir)
{
    Iter2D it{};
    it.start = r.start;
    it.end = r.end;
    if (r.end > r.start) {
        if (ir.row_range_length == 0) throw invalid_argument("row_range_length is 0 for a non-empty range");
        if (r.start < 0) throw invalid_argument("negative range start");
    }
    it.len = (int64_t) ir.row_range_length;
    it.first_row = it.len > 0 && r.end > r.start ? r.start / it.len : 0;
    it.last_row = it.len > 0 && r.end > r.start ? (r.end - 1) / it.len : -1;
    it.base = ir.row_range_start + ir.main_range_start;
    it.step = ir.main_range_step;
    return it;
}

Iter1D make_iter1d(const Range& r, const InnerRange1D& ir)
{
    if (r.end > r.start && r.start < 0) throw invalid_argument("negative range start");
    return Iter1D{r.start, r.end, ir.step, ir.start};
}

Span span_of(const Iter2D& it)
{
    Span s;
    if (it.end <= it.start) return s;
    auto idx = [&](int64_t row, int64_t col) { return it.base + row * it.step + col; };
    const int64_t c0 = it.start % it.len, c1 = (it.end - 1) % it.len;
    if (it.first_row == it.last_row) {
        s.add(idx(it.first_row, c0));
        s.add(idx(it.first_row, c1));
        return s;
    }
    s.add(idx(it.first_row, c0));
    s.add(idx(it.first_row, it.len - 1));
    s.add(idx(it.last_row, 0));
    s.add(idx(it.last_row, c1));
    if (it.last_row - it.first_row >= 2) {
        s.add(idx(it.first_row + 1, 0));
        s.add(idx(it.first_row + 1, it.len - 1));
        s.add(idx(it.last_row - 1, 0));
        s.add(idx(it.last_row - 1, it.len - 1));
    }
    return s;
}

Span span_of(const Iter1D& it)
{
    Span s;
    if (it.end <= it.start) return s;
    s.add(it.start * it.step + it.base);
    s.add((it.end - 1) * it.step + it.base);
    return s;
}

static bool ends_with(const char* str, const char* suffix)
{
    size_t n = strlen(str), m = strlen(suffix);
    return n >= m && strcmp(str + n - m, suffix) == 0;
}

void check_view(const view& v, const char* name, const Span& s, bool check_label)
{
    if (check_label && v.label && !ends_with(v.label, name))
        throw invalid_argument(std::string("Wrong view order: expected ") + name + ", got: " + v.label);
    if (s.empty()) return;
    if (!v.data) throw invalid_argument(std::string("null data pointer for view ") + name);
    if (s.lo < 0 || s.hi >= v.extent)
        throw invalid_argument(std::string("index out of bounds for view ") + name + ": [" + std::to_string(s.lo)
                               + ", " + std::to_string(s.hi) + "] not in [0, " + std::to_string(v.extent) + ")");
}

}  // namespace armon

// ---------------------------------------------------------------------------------------
// Iteration on the device
// ---------------------------------------------------------------------------------------

namespace {

constexpr int kBlock = 256;

// Visits every index of the range once. Each row's columns are shifted so that the waves'
// 64-element pieces of `ref` (the array the kernel streams) start on 128-byte lines: a row of
// the reference layout starts anywhere in a line (its ghost cells come first), and unaligned
// pieces touch 5 lines instead of 4 (measured on 7-stream kernels: 8% slower). The grid has
// kAlignSlack more columns than the row (grid_2d).
constexpr int kAlignSlack = 15;

template <class F>
__device__ __forceinline__ void for_each_2d(const Iter2D& it, const void* ref, F&& f)
{
    for (int64_t row = it.first_row + blockIdx.y; row <= it.last_row; row += gridDim.y) {
        const int64_t row0 = it.base + row * it.step;    // index of column 0
        const int64_t shift = (int64_t) (((uintptr_t) ref / sizeof(flt_t) + (uint64_t) row0) & kAlignSlack);
        const int64_t col = (int64_t) blockIdx.x * blockDim.x + threadIdx.x - shift;
        if (col < 0 || col >= it.len) continue;
        const int64_t k = row * it.len + col;
        if (k < it.start || k >= it.end) continue;
        f(row0 + col);
    }
}

template <class F>
__device__ __forceinline__ void for_each_1d(const Iter1D& it, F&& f)
{
    for (int64_t k = it.start + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < it.end;
         k += (int64_t) gridDim.x * blockDim.x)
        f(k, k * it.step + it.base);
}

dim3 grid_2d(const Iter2D& it)
{
    const int64_t rows = it.last_row - it.first_row + 1;
    const int64_t gx = (it.len + kBlock - 1) / kBlock;
    return dim3((unsigned) gx, (unsigned) (rows < 65535 ? rows : 65535), 1);
}

// The grid of for_each_2d: kAlignSlack columns more than the row.
dim3 grid_2d_aligned(const Iter2D& it)
{
    const int64_t rows = it.last_row - it.first_row + 1;
    const int64_t gx = (it.len + kAlignSlack + kBlock - 1) / kBlock;
    return dim3((unsigned) gx, (unsigned) (rows < 65535 ? rows : 65535), 1);
}

dim3 grid_1d(const Iter1D& it)
{
    const int64_t n = it.end - it.start;
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    return dim3((unsigned) (g > 0 ? g : 1), 1, 1);
}

bool empty_2d(const Iter2D& it) { return it.end <= it.start; }
bool empty_1d(const Iter1D& it) { return it.end <= it.start; }

void after_launch(const char* what) { hip_check(hipGetLastError(), what); }

// ---- kernels --------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_acoustic(Iter2D it, int64_t s, flt_t* __restrict__ ustar,
                                                     flt_t* __restrict__ pstar, const flt_t* __restrict__ rho,
                                                     const flt_t* __restrict__ u, const flt_t* __restrict__ pmat,
                                                     const flt_t* __restrict__ cmat)
{
    for_each_2d(it, rho, [&](int64_t i) {
        flt_t us, ps;
        godunov<flt_t>(rho[i], rho[i - s], cmat[i], cmat[i - s], u[i], u[i - s], pmat[i], pmat[i - s], us, ps);
        ustar[i] = us;
        pstar[i] = ps;
    });
}

template <int L>
__global__ void __launch_bounds__(kBlock) k_acoustic_GAD(Iter2D it, int64_t s, flt_t dt, flt_t dx,
                                                         flt_t* __restrict__ ustar, flt_t* __restrict__ pstar,
                                                         const flt_t* __restrict__ rho, const flt_t* __restrict__ u,
                                                         const flt_t* __restrict__ pmat, const flt_t* __restrict__ cmat)
{
    for_each_2d(it, rho, [&](int64_t i) {
        flt_t us_im, ps_im, us_i, ps_i, us_ip, ps_ip;
        godunov<flt_t>(rho[i - s], rho[i - 2 * s], cmat[i - s], cmat[i - 2 * s], u[i - s], u[i - 2 * s],
                       pmat[i - s], pmat[i - 2 * s], us_im, ps_im);
        godunov<flt_t>(rho[i], rho[i - s], cmat[i], cmat[i - s], u[i], u[i - s], pmat[i], pmat[i - s], us_i, ps_i);
        godunov<flt_t>(rho[i + s], rho[i], cmat[i + s], cmat[i], u[i + s], u[i], pmat[i + s], pmat[i], us_ip, ps_ip);
        flt_t us, ps;
        gad_correct<L, flt_t>(dt, dx, rho[i - s], rho[i], cmat[i - s], cmat[i], u[i - s], u[i], pmat[i - s], pmat[i],
                              us_im, ps_im, us_i, ps_i, us_ip, ps_ip, us, ps);
        ustar[i] = us;
        pstar[i] = ps;
    });
}

// The GAD flux of acoustic.cpp:40-88 solves the Godunov problem at faces i-s, i and i+s for
// every face i, so each face's solution is computed three times. The two kernels below solve
// it once per face and hand it to the neighbouring faces (same arithmetic, same bits):
//  - along rows (s = 1), through LDS: a block covers kGadTile consecutive faces of one row and
//    writes the kGadTile - 2 inner ones;
//  - across rows (s = the row step), in registers: each thread marches down a run of rows of
//    its column with the last three solutions in a sliding window.
// A face's solution is computed only where an active face (one of the range) or one of its
// two neighbours needs it, so the kernels read nothing the generic kernel would not.
constexpr int kGadTile = 256;
constexpr int kGadOut = kGadTile - 2;

// Active columns [lo, hi) of `row` of the range (a partial first or last row).
__device__ __forceinline__ void row_columns(const Iter2D& it, int64_t row, int64_t& lo, int64_t& hi)
{
    lo = row == it.first_row ? it.start - row * it.len : 0;
    hi = row == it.last_row ? it.end - row * it.len : it.len;
}

template <int L>
__global__ void __launch_bounds__(kGadTile) k_acoustic_GAD_x(Iter2D it, flt_t dt, flt_t dx,
                                                            flt_t* __restrict__ ustar, flt_t* __restrict__ pstar,
                                                            const flt_t* __restrict__ rho,
                                                            const flt_t* __restrict__ u,
                                                            const flt_t* __restrict__ pmat,
                                                            const flt_t* __restrict__ cmat)
{
    __shared__ flt_t g_u[kGadTile], g_p[kGadTile];
    const int t = threadIdx.x;
    const int64_t col = (int64_t) blockIdx.x * kGadOut - 1 + t;   // this thread's face
    for (int64_t row = it.first_row + blockIdx.y; row <= it.last_row; row += gridDim.y) {
        int64_t lo, hi;
        row_columns(it, row, lo, hi);
        const int64_t i = it.base + row * it.step + col;
        flt_t r_c = 0, r_m = 0, c_c = 0, c_m = 0, u_c = 0, u_m = 0, p_c = 0, p_m = 0;
        if (col >= lo - 1 && col <= hi) {
            r_c = rho[i]; r_m = rho[i - 1];
            c_c = cmat[i]; c_m = cmat[i - 1];
            u_c = u[i]; u_m = u[i - 1];
            p_c = pmat[i]; p_m = pmat[i - 1];
            godunov<flt_t>(r_c, r_m, c_c, c_m, u_c, u_m, p_c, p_m, g_u[t], g_p[t]);
        }
        __syncthreads();
        if (t >= 1 && t <= kGadOut && col >= lo && col < hi) {
            flt_t us, ps;
            gad_correct<L, flt_t>(dt, dx, r_m, r_c, c_m, c_c, u_m, u_c, p_m, p_c, g_u[t - 1], g_p[t - 1], g_u[t],
                                  g_p[t], g_u[t + 1], g_p[t + 1], us, ps);
            ustar[i] = us;
            pstar[i] = ps;
        }
        __syncthreads();   // the next row reuses the LDS
    }
}

template <int L>
__global__ void __launch_bounds__(kBlock) k_acoustic_GAD_y(Iter2D it, int64_t run, flt_t dt, flt_t dx,
                                                           flt_t* __restrict__ ustar, flt_t* __restrict__ pstar,
                                                           const flt_t* __restrict__ rho, const flt_t* __restrict__ u,
                                                           const flt_t* __restrict__ pmat,
                                                           const flt_t* __restrict__ cmat)
{
    const int64_t col = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= it.len) return;
    // rows of this column in the range: [r_lo, r_hi]
    const int64_t r_lo = it.first_row + (col < it.start - it.first_row * it.len ? 1 : 0);
    const int64_t r_hi = it.last_row - (col > it.end - 1 - it.last_row * it.len ? 1 : 0);
    const int64_t s = it.step;
    const int64_t nruns = (it.last_row - it.first_row + run) / run;
    for (int64_t k = blockIdx.y; k < nruns; k += gridDim.y) {
        const int64_t ra = max(it.first_row + k * run, r_lo);
        const int64_t rb = min(it.first_row + (k + 1) * run - 1, r_hi);
        if (ra > rb) continue;
        // cells of rows ra - 2 (mm), ra - 1 (m), and the solutions at faces ra - 1 and ra
        int64_t i = it.base + ra * s + col;
        flt_t r_mm = rho[i - 2 * s], c_mm = cmat[i - 2 * s], u_mm = u[i - 2 * s], p_mm = pmat[i - 2 * s];
        flt_t r_m = rho[i - s], c_m = cmat[i - s], u_m = u[i - s], p_m = pmat[i - s];
        flt_t r_c = rho[i], c_c = cmat[i], u_c = u[i], p_c = pmat[i];
        flt_t gu_m, gp_m, gu_c, gp_c;
        godunov<flt_t>(r_m, r_mm, c_m, c_mm, u_m, u_mm, p_m, p_mm, gu_m, gp_m);
        godunov<flt_t>(r_c, r_m, c_c, c_m, u_c, u_m, p_c, p_m, gu_c, gp_c);
        for (int64_t row = ra; row <= rb; row++, i += s) {
            const flt_t r_p = rho[i + s], c_p = cmat[i + s], u_p = u[i + s], p_p = pmat[i + s];
            flt_t gu_p, gp_p;
            godunov<flt_t>(r_p, r_c, c_p, c_c, u_p, u_c, p_p, p_c, gu_p, gp_p);
            flt_t us, ps;
            gad_correct<L, flt_t>(dt, dx, r_m, r_c, c_m, c_c, u_m, u_c, p_m, p_c, gu_m, gp_m, gu_c, gp_c, gu_p, gp_p,
                                  us, ps);
            ustar[i] = us;
            pstar[i] = ps;
            r_m = r_c; c_m = c_c; u_m = u_c; p_m = p_c;
            r_c = r_p; c_c = c_p; u_c = u_p; p_c = p_p;
            gu_m = gu_c; gp_m = gp_c;
            gu_c = gu_p; gp_c = gp_p;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_perfect_gas_EOS(Iter2D it, flt_t gamma, const flt_t* __restrict__ rho,
                                                            const flt_t* __restrict__ Emat,
                                                            const flt_t* __restrict__ umat,
                                                            const flt_t* __restrict__ vmat, flt_t* __restrict__ pmat,
                                                            flt_t* __restrict__ cmat, flt_t* __restrict__ gmat)
{
    for_each_2d(it, rho, [&](int64_t i) {
        flt_t p, c, g;
        perfect_gas<flt_t>(gamma, rho[i], Emat[i], umat[i], vmat[i], p, c, g);
        pmat[i] = p;
        cmat[i] = c;
        gmat[i] = g;
    });
}

__global__ void __launch_bounds__(kBlock) k_bizarrium_EOS(Iter2D it, const flt_t* __restrict__ rho,
                                                          const flt_t* __restrict__ umat, const flt_t* __restrict__ vmat,
                                                          const flt_t* __restrict__ Emat, flt_t* __restrict__ pmat,
                                                          flt_t* __restrict__ cmat, flt_t* __restrict__ gmat)
{
    for_each_2d(it, rho, [&](int64_t i) {
        double p, c, g;
        bizarrium((double) rho[i], (double) Emat[i], (double) umat[i], (double) vmat[i], p, c, g);
        pmat[i] = (flt_t) p;
        cmat[i] = (flt_t) c;
        gmat[i] = (flt_t) g;
    });
}

__global__ void __launch_bounds__(kBlock) k_cell_update(Iter2D it, int64_t s, flt_t dx, flt_t dt,
                                                        const flt_t* __restrict__ ustar,
                                                        const flt_t* __restrict__ pstar, flt_t* __restrict__ rho,
                                                        flt_t* __restrict__ u, flt_t* __restrict__ Emat)
{
    for_each_2d(it, rho, [&](int64_t i) {
        flt_t r = rho[i], uu = u[i], E = Emat[i];
        armon::cell_update<flt_t>(dx, dt, ustar[i], ustar[i + s], pstar[i], pstar[i + s], r, uu, E);
        rho[i] = r;
        u[i] = uu;
        Emat[i] = E;
    });
}

// DXM: how x / dx is computed (DxMode), with inv_dx = 1 / dx for DX_POW2 and DX_RECIP; every mode
// gives the IEEE quotient's bits.
template <int DXM>
__global__ void __launch_bounds__(kBlock) k_euler_projection(Iter2D it, int64_t s, flt_t dx, flt_t inv_dx, flt_t dt,
                                                             const flt_t* __restrict__ ustar, flt_t* __restrict__ rho,
                                                             flt_t* __restrict__ umat, flt_t* __restrict__ vmat,
                                                             flt_t* __restrict__ Emat, const flt_t* __restrict__ a_rho,
                                                             const flt_t* __restrict__ a_urho,
                                                             const flt_t* __restrict__ a_vrho,
                                                             const flt_t* __restrict__ a_Erho)
{
    for_each_2d(it, rho, [&](int64_t i) {
        flt_t r = rho[i], u = umat[i], v = vmat[i], E = Emat[i];
        projection<DXM, flt_t>(dx, inv_dx, dt, ustar[i], ustar[i + s], a_rho[i], a_rho[i + s], a_urho[i],
                               a_urho[i + s], a_vrho[i], a_vrho[i + s], a_Erho[i], a_Erho[i + s], r, u, v, E);
        rho[i] = r;
        umat[i] = u;
        vmat[i] = v;
        Emat[i] = E;
    });
}

__global__ void __launch_bounds__(kBlock) k_advection_first_order(Iter2D it, int64_t s, flt_t dt,
                                                                  const flt_t* __restrict__ ustar,
                                                                  const flt_t* __restrict__ rho,
                                                                  const flt_t* __restrict__ umat,
                                                                  const flt_t* __restrict__ vmat,
                                                                  const flt_t* __restrict__ Emat,
                                                                  flt_t* __restrict__ a_rho, flt_t* __restrict__ a_urho,
                                                                  flt_t* __restrict__ a_vrho, flt_t* __restrict__ a_Erho)
{
    for_each_2d(it, ustar, [&](int64_t is) {
        flt_t disp = dt * ustar[is];
        const int64_t i = is - (int64_t) (disp > 0) * s;
        flt_t f0, f1, f2, f3;
        advection_flux_1<flt_t>(disp, rho[i], umat[i], vmat[i], Emat[i], f0, f1, f2, f3);
        a_rho[is] = f0;
        a_urho[is] = f1;
        a_vrho[is] = f2;
        a_Erho[is] = f3;
    });
}

__global__ void __launch_bounds__(kBlock) k_advection_second_order(Iter2D it, int64_t s, flt_t dx, flt_t dt,
                                                                   const flt_t* __restrict__ ustar,
                                                                   const flt_t* __restrict__ rho,
                                                                   const flt_t* __restrict__ umat,
                                                                   const flt_t* __restrict__ vmat,
                                                                   const flt_t* __restrict__ Emat,
                                                                   flt_t* __restrict__ a_rho,
                                                                   flt_t* __restrict__ a_urho,
                                                                   flt_t* __restrict__ a_vrho,
                                                                   flt_t* __restrict__ a_Erho)
{
    for_each_2d(it, ustar, [&](int64_t is) {
        flt_t disp = dt * ustar[is];
        flt_t Dx;
        int64_t i = is;
        if (disp > 0) {
            Dx = -(dx - dt * ustar[is - s]);
            i = is - s;
        } else {
            Dx = dx + dt * ustar[is + s];
        }
        flt_t f0, f1, f2, f3;
        advection_flux_2<flt_t>(dx, dt, disp, Dx, ustar[i - s], ustar[i], ustar[i + s], ustar[i + 2 * s], rho[i - s],
                                rho[i], rho[i + s], umat[i - s], umat[i], umat[i + s], vmat[i - s], vmat[i],
                                vmat[i + s], Emat[i - s], Emat[i], Emat[i + s], f0, f1, f2, f3);
        a_rho[is] = f0;
        a_urho[is] = f1;
        a_vrho[is] = f2;
        a_Erho[is] = f3;
    });
}

// advection_second_order (euler_projection.cpp:83-118) reads, for each face i, the face velocities
// u* at i-2s..i+2s and the state at i-2s..i+s (the donor side's part of them) straight from
// memory, 18 loads per face. The two kernels below load each element once and keep the
// neighbours close (same arithmetic, same bits):
//  - along rows (s = 1), through LDS: a block stages kAdvTile consecutive elements of one row
//    and writes the kAdvTile - 4 inner faces;
//  - across rows (s = the row step), in registers: each thread marches down a run of rows of
//    its column with windows of the last rows.
// Each loads only elements of the spans that advection_second_order checks.
constexpr int kAdvTile = 256;
constexpr int kAdvOut = kAdvTile - 4;

__global__ void __launch_bounds__(kAdvTile) k_advection_second_order_x(Iter2D it, flt_t dx, flt_t dt,
                                                                      const flt_t* __restrict__ ustar,
                                                                      const flt_t* __restrict__ rho,
                                                                      const flt_t* __restrict__ umat,
                                                                      const flt_t* __restrict__ vmat,
                                                                      const flt_t* __restrict__ Emat,
                                                                      flt_t* __restrict__ a_rho,
                                                                      flt_t* __restrict__ a_urho,
                                                                      flt_t* __restrict__ a_vrho,
                                                                      flt_t* __restrict__ a_Erho)
{
    __shared__ flt_t s_us[kAdvTile], s_r[kAdvTile], s_u[kAdvTile], s_v[kAdvTile], s_E[kAdvTile];
    const int t = threadIdx.x;
    const int64_t col = (int64_t) blockIdx.x * kAdvOut - 2 + t;   // this thread's element
    for (int64_t row = it.first_row + blockIdx.y; row <= it.last_row; row += gridDim.y) {
        int64_t lo, hi;
        row_columns(it, row, lo, hi);
        const int64_t i = it.base + row * it.step + col;
        // u* of faces lo-2 .. hi+1, the state of cells lo-2 .. hi
        if (col >= lo - 2 && col <= hi + 1) s_us[t] = ustar[i];
        if (col >= lo - 2 && col <= hi) {
            s_r[t] = rho[i];
            s_u[t] = umat[i];
            s_v[t] = vmat[i];
            s_E[t] = Emat[i];
        }
        __syncthreads();
        if (t >= 2 && t < 2 + kAdvOut && col >= lo && col < hi) {
            const flt_t disp = dt * s_us[t];
            flt_t Dx;
            int d = t;
            if (disp > 0) {
                Dx = -(dx - dt * s_us[t - 1]);
                d = t - 1;
            } else {
                Dx = dx + dt * s_us[t + 1];
            }
            flt_t f0, f1, f2, f3;
            advection_flux_2<flt_t>(dx, dt, disp, Dx, s_us[d - 1], s_us[d], s_us[d + 1], s_us[d + 2], s_r[d - 1],
                                    s_r[d], s_r[d + 1], s_u[d - 1], s_u[d], s_u[d + 1], s_v[d - 1], s_v[d], s_v[d + 1],
                                    s_E[d - 1], s_E[d], s_E[d + 1], f0, f1, f2, f3);
            a_rho[i] = f0;
            a_urho[i] = f1;
            a_vrho[i] = f2;
            a_Erho[i] = f3;
        }
        __syncthreads();   // the next row reuses the LDS
    }
}

__global__ void __launch_bounds__(kBlock) k_advection_second_order_y(Iter2D it, int64_t run, flt_t dx, flt_t dt,
                                                                    const flt_t* __restrict__ ustar,
                                                                    const flt_t* __restrict__ rho,
                                                                    const flt_t* __restrict__ umat,
                                                                    const flt_t* __restrict__ vmat,
                                                                    const flt_t* __restrict__ Emat,
                                                                    flt_t* __restrict__ a_rho,
                                                                    flt_t* __restrict__ a_urho,
                                                                    flt_t* __restrict__ a_vrho,
                                                                    flt_t* __restrict__ a_Erho)
{
    const int64_t col = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= it.len) return;
    // rows of this column in the range: [r_lo, r_hi]
    const int64_t r_lo = it.first_row + (col < it.start - it.first_row * it.len ? 1 : 0);
    const int64_t r_hi = it.last_row - (col > it.end - 1 - it.last_row * it.len ? 1 : 0);
    const int64_t s = it.step;
    const int64_t nruns = (it.last_row - it.first_row + run) / run;
    for (int64_t k = blockIdx.y; k < nruns; k += gridDim.y) {
        const int64_t ra = max(it.first_row + k * run, r_lo);
        const int64_t rb = min(it.first_row + (k + 1) * run - 1, r_hi);
        if (ra > rb) continue;
        int64_t i = it.base + ra * s + col;
        // windows at face/cell row `row`: u* of rows row-2 .. row+2, the state of rows row-2 .. row+1
        flt_t w_us[5], w_r[4], w_u[4], w_v[4], w_E[4];
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) {
            const int64_t j = i + (k2 - 2) * s;
            w_us[k2] = ustar[j];
            w_r[k2] = rho[j];
            w_u[k2] = umat[j];
            w_v[k2] = vmat[j];
            w_E[k2] = Emat[j];
        }
        w_us[4] = ustar[i + 2 * s];
        for (int64_t row = ra; row <= rb; row++, i += s) {
            const flt_t disp = dt * w_us[2];
            flt_t f0, f1, f2, f3;
            if (disp > 0) {
                const flt_t Dx = -(dx - dt * w_us[1]);
                // donor row-1: u* of rows row-2 .. row+1, the state of rows row-2 .. row
                advection_flux_2<flt_t>(dx, dt, disp, Dx, w_us[0], w_us[1], w_us[2], w_us[3], w_r[0], w_r[1], w_r[2],
                                        w_u[0], w_u[1], w_u[2], w_v[0], w_v[1], w_v[2], w_E[0], w_E[1], w_E[2], f0, f1,
                                        f2, f3);
            } else {
                const flt_t Dx = dx + dt * w_us[3];
                // donor row: u* of rows row-1 .. row+2, the state of rows row-1 .. row+1
                advection_flux_2<flt_t>(dx, dt, disp, Dx, w_us[1], w_us[2], w_us[3], w_us[4], w_r[1], w_r[2], w_r[3],
                                        w_u[1], w_u[2], w_u[3], w_v[1], w_v[2], w_v[3], w_E[1], w_E[2], w_E[3], f0, f1,
                                        f2, f3);
            }
            a_rho[i] = f0;
            a_urho[i] = f1;
            a_vrho[i] = f2;
            a_Erho[i] = f3;
            if (row < rb) {
#pragma unroll
                for (int k2 = 0; k2 < 4; k2++) w_us[k2] = w_us[k2 + 1];
#pragma unroll
                for (int k2 = 0; k2 < 3; k2++) {
                    w_r[k2] = w_r[k2 + 1];
                    w_u[k2] = w_u[k2 + 1];
                    w_v[k2] = w_v[k2 + 1];
                    w_E[k2] = w_E[k2 + 1];
                }
                w_us[4] = ustar[i + 3 * s];
                w_r[3] = rho[i + 2 * s];
                w_u[3] = umat[i + 2 * s];
                w_v[3] = vmat[i + 2 * s];
                w_E[3] = Emat[i + 2 * s];
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_boundary_conditions(Iter1D it, int64_t stencil_width, int64_t disp,
                                                                flt_t u_factor, flt_t v_factor, flt_t* rho,
                                                                flt_t* umat, flt_t* vmat, flt_t* pmat, flt_t* cmat,
                                                                flt_t* gmat, flt_t* Emat)
{
    for_each_1d(it, [&](int64_t, int64_t i) {
        int64_t ip = i + disp;
        for (int64_t w = 0; w < stencil_width; w++) {
            rho[i] = rho[ip];
            umat[i] = umat[ip] * u_factor;
            vmat[i] = vmat[ip] * v_factor;
            pmat[i] = pmat[ip];
            cmat[i] = cmat[ip];
            gmat[i] = gmat[ip];
            Emat[i] = Emat[ip];
            i -= disp;
            ip += disp;
        }
    });
}

__global__ void __launch_bounds__(kBlock) k_read_border(Iter2D it, int64_t side_length, int64_t nghost,
                                                        const flt_t* __restrict__ rho, const flt_t* __restrict__ umat,
                                                        const flt_t* __restrict__ vmat, const flt_t* __restrict__ pmat,
                                                        const flt_t* __restrict__ cmat, const flt_t* __restrict__ gmat,
                                                        const flt_t* __restrict__ Emat, flt_t* __restrict__ values)
{
    const int64_t col = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= it.len) return;
    for (int64_t row = it.first_row + blockIdx.y; row <= it.last_row; row += gridDim.y) {
        const int64_t k = row * it.len + col;
        if (k < it.start || k >= it.end) continue;
        const int64_t idx = it.base + row * it.step + col;
        const int64_t a = ((k % nghost) * side_length + k / nghost) * 7;
        values[a + 0] = rho[idx];
        values[a + 1] = umat[idx];
        values[a + 2] = vmat[idx];
        values[a + 3] = pmat[idx];
        values[a + 4] = cmat[idx];
        values[a + 5] = gmat[idx];
        values[a + 6] = Emat[idx];
    }
}

__global__ void __launch_bounds__(kBlock) k_write_border(Iter2D it, int64_t side_length, int64_t nghost,
                                                         flt_t* __restrict__ rho, flt_t* __restrict__ umat,
                                                         flt_t* __restrict__ vmat, flt_t* __restrict__ pmat,
                                                         flt_t* __restrict__ cmat, flt_t* __restrict__ gmat,
                                                         flt_t* __restrict__ Emat, const flt_t* __restrict__ values)
{
    const int64_t col = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= it.len) return;
    for (int64_t row = it.first_row + blockIdx.y; row <= it.last_row; row += gridDim.y) {
        const int64_t k = row * it.len + col;
        if (k < it.start || k >= it.end) continue;
        const int64_t idx = it.base + row * it.step + col;
        const int64_t a = ((k % nghost) * side_length + k / nghost) * 7;
        rho[idx] = values[a + 0];
        umat[idx] = values[a + 1];
        vmat[idx] = values[a + 2];
        pmat[idx] = values[a + 3];
        cmat[idx] = values[a + 4];
        gmat[idx] = values[a + 5];
        Emat[idx] = values[a + 6];
    }
}

// init_test.cpp:12-48 (states) and test_cases.h:50-116 (regions)
struct TestState {
    flt_t high_rho, low_rho, high_E, low_E, high_u, low_u, high_v, low_v;
};

TestState test_state(int test, flt_t r)
{
    TestState t{};
    switch (test) {
    case 3:   // Bizarrium
        t.high_rho = flt_t(1.42857142857e+4); t.low_rho = flt_t(10000.);
        t.high_E = flt_t(4.48657821135e+6);   t.low_E = flt_t(31250);
        t.high_u = 0;                         t.low_u = flt_t(250.);
        break;
    case 4:   // Sedov
        t.high_rho = 1; t.low_rho = 1;
        t.high_E = flt_t(0.851072 / (M_PI * (double) (r * r)));
        t.low_E = flt_t(2.5e-14);
        break;
    default:  // Sod, Sod_y, Sod_circ
        t.high_rho = 1;            t.low_rho = flt_t(0.125);
        t.high_E = flt_t(2.5);     t.low_E = flt_t(2.0);
        break;
    }
    return t;
}

__device__ __forceinline__ bool region_high(int test, flt_t r, flt_t x, flt_t y)
{
    switch (test) {
    case 0: return x <= flt_t(0.5);
    case 1: return y <= flt_t(0.5);
    case 2: return ((x - flt_t(0.5)) * (x - flt_t(0.5)) + (y - flt_t(0.5)) * (y - flt_t(0.5))) <= flt_t(0.09);
    case 3: return x <= flt_t(0.5);
    default: return (x * x + y * y) <= r * r;
    }
}

struct InitArgs {
    int64_t row_length, nb_ghosts, nx, ny, pos_x, pos_y, g_nx, g_ny;
    flt_t sx, sy, ox, oy, r;
    int test;
    bool debug_indexes;
    TestState st;
};

__global__ void __launch_bounds__(kBlock) k_init_test(Iter2D it, InitArgs a, flt_t* x, flt_t* y, flt_t* rho,
                                                      flt_t* Emat, flt_t* umat, flt_t* vmat, flt_t* mask, flt_t* pmat,
                                                      flt_t* cmat, flt_t* gmat, flt_t* ustar, flt_t* pstar)
{
    for_each_2d(it, rho, [&](int64_t i) {
        const int64_t ix = (i % a.row_length) - a.nb_ghosts;
        const int64_t iy = (i / a.row_length) - a.nb_ghosts;
        const int64_t g_ix = ix + a.pos_x;
        const int64_t g_iy = iy + a.pos_y;
        const flt_t xi = flt_t(g_ix) / flt_t(a.nx) * a.sx + a.ox;
        const flt_t yi = flt_t(g_iy) / flt_t(a.ny) * a.sy + a.oy;
        x[i] = xi;
        y[i] = yi;
        const flt_t x_mid = xi + a.sx / flt_t(2 * a.g_nx);
        const flt_t y_mid = yi + a.sy / flt_t(2 * a.g_ny);
        if (a.debug_indexes) {
            rho[i] = flt_t(i + 1);
            Emat[i] = flt_t(i + 1);
            umat[i] = flt_t(i + 1);
            vmat[i] = flt_t(i + 1);
        } else if (region_high(a.test, a.r, x_mid, y_mid)) {
            rho[i] = a.st.high_rho;
            Emat[i] = a.st.high_E;
            umat[i] = a.st.high_u;
            vmat[i] = a.st.high_v;
        } else {
            rho[i] = a.st.low_rho;
            Emat[i] = a.st.low_E;
            umat[i] = a.st.low_u;
            vmat[i] = a.st.low_v;
        }
        mask[i] = (0 <= ix && ix < a.nx && 0 <= iy && iy < a.ny) ? flt_t(1) : flt_t(0);
        pmat[i] = 0;
        cmat[i] = 1;
        gmat[i] = 0;
        ustar[i] = 0;
        pstar[i] = 0;
    });
}

// ---- reductions -------------------------------------------------------------------------

constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = kmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Min of the CFL bound over the range. The per-thread fold starts from the largest finite
// value, Kokkos' identity for a Min reduction (so a range where every cell gives +inf
// returns DBL_MAX, like the reference). The block minima are merged with an integer
// atomicMin on the bit pattern, which orders non-negative doubles (NaN bounds rank above
// +inf and never win, as with the reference's comparison-based min).
__global__ void __launch_bounds__(kBlock) k_dt_CFL(Iter1D it, flt_t dx, flt_t dy, const flt_t* __restrict__ umat,
                                                   const flt_t* __restrict__ vmat, const flt_t* __restrict__ cmat,
                                                   const flt_t* __restrict__ mask, unsigned long long* min_bits)
{
    flt_t local = (sizeof(flt_t) == 8) ? flt_t(DBL_MAX) : flt_t(FLT_MAX);
    for_each_1d(it, [&](int64_t, int64_t i) {
        local = kmin(local, cfl_cell<flt_t>(dx, dy, umat[i], vmat[i], cmat[i], mask[i]));
    });
    double v = (double) local;
    v = wave_min(v);
    __shared__ double smem[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = smem[0];
        for (int w = 1; w < kBlock / 64; w++) m = kmin(m, smem[w]);
        atomicMin(min_bits, (unsigned long long) __double_as_longlong(m));
    }
}

__global__ void k_fill_bits(unsigned long long* p, unsigned long long v) { *p = v; }

__global__ void k_bits_to_double(const unsigned long long* bits, double* out)
{
    *out = __longlong_as_double((long long) *bits);
}

// Compensated (Neumaier) sums of mass and energy; per-block partial (sum, compensation)
// pairs, combined in a fixed order by k_conservation_final: run-to-run deterministic.
struct Neumaier {
    double s = 0, c = 0;
    __device__ __forceinline__ void add(double v)
    {
        double t = s + v;
        if (abs_t<double>(s) >= abs_t<double>(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
};

__global__ void __launch_bounds__(kBlock) k_conservation_partial(Iter1D it, double ds, const flt_t* __restrict__ rho,
                                                                 const flt_t* __restrict__ Emat,
                                                                 const flt_t* __restrict__ mask, double* partials)
{
    Neumaier m, e;
    for_each_1d(it, [&](int64_t, int64_t i) {
        double cell_mass = (double) rho[i] * (double) mask[i] * ds;
        double cell_energy = cell_mass * (double) Emat[i];
        m.add(cell_mass);
        e.add(cell_energy);
    });
    __shared__ double sm[4][kBlock];
    sm[0][threadIdx.x] = m.s;
    sm[1][threadIdx.x] = m.c;
    sm[2][threadIdx.x] = e.s;
    sm[3][threadIdx.x] = e.c;
    __syncthreads();
    if (threadIdx.x == 0) {
        Neumaier bm, be;
        for (int t = 0; t < kBlock; t++) {
            bm.add(sm[0][t]); bm.add(sm[1][t]);
            be.add(sm[2][t]); be.add(sm[3][t]);
        }
        partials[4 * blockIdx.x + 0] = bm.s;
        partials[4 * blockIdx.x + 1] = bm.c;
        partials[4 * blockIdx.x + 2] = be.s;
        partials[4 * blockIdx.x + 3] = be.c;
    }
}

__global__ void k_conservation_final(int nblocks, const double* partials, double* out)
{
    Neumaier m, e;
    for (int b = 0; b < nblocks; b++) {
        m.add(partials[4 * b + 0]); m.add(partials[4 * b + 1]);
        e.add(partials[4 * b + 2]); e.add(partials[4 * b + 3]);
    }
    out[0] = m.s + m.c;
    out[1] = e.s + e.c;
}

// ---- allocation-free conservation sums ----------------------------------------------------

// Scratch of conservation_vars_noalloc: 4 doubles of partials per block (at most kRedBlocks
// blocks) and the 2 results, 32,784 bytes, held by every device the library runs on until the
// process ends. Host code never uses this name as an address: the device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch below).
constexpr int kNoAllocBlocks = kRedBlocks;
constexpr int kNoAllocDoubles = 4 * kNoAllocBlocks + 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;

// The device address of `scratch` on the current device, looked up on the first call on that
// device (with it current) and kept.
double* noalloc_scratch()
{
    static std::mutex mtx;
    static double* ptr[kMaxDevices] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(mtx);
    if (!ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

const char* limiter_error(int limiter)
{
    static thread_local char buf[160];
    snprintf(buf, sizeof buf, "Invalid limiter index: %d, expected a value between 0 and 2", limiter);
    return buf;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------------------

extern "C" {

ARMON_EXPORT void acoustic(const Range& range, const InnerRange2D& inner_range, Idx s, view& ustar, view& pstar,
                           const view& rho, const view& u, const view& pmat, const view& cmat)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    const Span sp_n = sp.merged(sp.shifted(-s, -s));
    check_view(ustar, "ustar", sp);
    check_view(pstar, "pstar", sp);
    check_view(rho, "rho", sp_n);
    check_view(u, "u", sp_n, false);
    check_view(pmat, "pmat", sp_n);
    check_view(cmat, "cmat", sp_n);
    if (empty_2d(it)) return;
    k_acoustic<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, ustar.data, pstar.data, rho.data, u.data, pmat.data,
                                                      cmat.data);
    after_launch("acoustic");
}
ARMON_KERNEL_CATCH("acoustic")

ARMON_EXPORT void acoustic_GAD(const Range& range, const InnerRange2D& inner_range, Idx s, flt_t dt, flt_t dx,
                               view& ustar, view& pstar, const view& rho, const view& u, const view& pmat,
                               const view& cmat, Limiter limiter)
ARMON_KERNEL_TRY
{
    const int L = static_cast<int>(limiter);
    if (L < 0 || L > 2) throw out_of_range(limiter_error(L));
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    const Span sp_n = sp.merged(sp.shifted(-2 * s, -2 * s)).merged(sp.shifted(s, s));
    check_view(ustar, "ustar", sp);
    check_view(pstar, "pstar", sp);
    check_view(rho, "rho", sp_n);
    check_view(u, "u", sp_n, false);
    check_view(pmat, "pmat", sp_n);
    check_view(cmat, "cmat", sp_n);
    if (empty_2d(it)) return;
    const int64_t rows = it.last_row - it.first_row + 1;
    auto launch = [&]<int LL>() {
        if (s == 1) {
            // one block per kGadOut columns of a row
            const dim3 g((unsigned) ((it.len + kGadOut - 1) / kGadOut), (unsigned) (rows < 65535 ? rows : 65535));
            k_acoustic_GAD_x<LL><<<g, kGadTile, 0, stream()>>>(it, dt, dx, ustar.data, pstar.data, rho.data, u.data,
                                                               pmat.data, cmat.data);
        } else if (s == it.step) {
            // runs of 8 rows (4 while that leaves fewer than 8192 waves of 64 columns): these
            // kernels wait on memory, and short runs keep more waves in flight; the rows a run
            // re-reads before its first output come from L2 (measured at 8192^2: 8 rows 12-16%
            // faster than 64, 4 rows the same as 8)
            int64_t run = 8;
            const int64_t groups = (it.len + 63) / 64;
            if (groups * ((rows + run - 1) / run) < 8192) run = 4;
            const int64_t nruns = (rows + run - 1) / run;
            const dim3 g((unsigned) ((it.len + kBlock - 1) / kBlock), (unsigned) (nruns < 65535 ? nruns : 65535));
            k_acoustic_GAD_y<LL><<<g, kBlock, 0, stream()>>>(it, run, dt, dx, ustar.data, pstar.data, rho.data,
                                                             u.data, pmat.data, cmat.data);
        } else {
            k_acoustic_GAD<LL><<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, dt, dx, ustar.data, pstar.data, rho.data,
                                                                      u.data, pmat.data, cmat.data);
        }
    };
    switch (L) {
    case 0: launch.template operator()<0>(); break;
    case 1: launch.template operator()<1>(); break;
    default: launch.template operator()<2>(); break;
    }
    after_launch("acoustic_GAD");
}
ARMON_KERNEL_CATCH("acoustic_GAD")

ARMON_EXPORT void perfect_gas_EOS(const Range& range, const InnerRange2D& inner_range, flt_t gamma, const view& rho,
                                  const view& Emat, const view& umat, const view& vmat, view& pmat, view& cmat,
                                  view& gmat)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    check_view(rho, "rho", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(Emat, "Emat", sp);
    check_view(cmat, "cmat", sp);
    check_view(pmat, "pmat", sp);
    check_view(gmat, "gmat", sp);
    if (empty_2d(it)) return;
    k_perfect_gas_EOS<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, gamma, rho.data, Emat.data, umat.data, vmat.data,
                                                             pmat.data, cmat.data, gmat.data);
    after_launch("perfect_gas_EOS");
}
ARMON_KERNEL_CATCH("perfect_gas_EOS")

ARMON_EXPORT void bizarrium_EOS(const Range& range, const InnerRange2D& inner_range, const view& rho,
                                const view& umat, const view& vmat, const view& Emat, view& pmat, view& cmat,
                                view& gmat)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    check_view(rho, "rho", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(Emat, "Emat", sp);
    check_view(cmat, "cmat", sp);
    check_view(pmat, "pmat", sp);
    check_view(gmat, "gmat", sp);
    if (empty_2d(it)) return;
    k_bizarrium_EOS<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, rho.data, umat.data, vmat.data, Emat.data, pmat.data,
                                                           cmat.data, gmat.data);
    after_launch("bizarrium_EOS");
}
ARMON_KERNEL_CATCH("bizarrium_EOS")

ARMON_EXPORT void cell_update(const Range& range, const InnerRange2D& inner_range, Idx s, flt_t dx, flt_t dt,
                              const view& ustar, const view& pstar, view& rho, view& u, view& Emat)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    const Span sp_f = sp.merged(sp.shifted(s, s));
    check_view(ustar, "ustar", sp_f);
    check_view(pstar, "pstar", sp_f);
    check_view(Emat, "Emat", sp);
    check_view(rho, "rho", sp);
    check_view(u, "u", sp, false);
    if (empty_2d(it)) return;
    k_cell_update<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, dx, dt, ustar.data, pstar.data, rho.data, u.data,
                                                         Emat.data);
    after_launch("cell_update");
}
ARMON_KERNEL_CATCH("cell_update")

ARMON_EXPORT void euler_projection(const Range& range, const InnerRange2D& inner_range, Idx s, flt_t dx, flt_t dt,
                                   const view& ustar, view& rho, view& umat, view& vmat, view& Emat,
                                   const view& advection_rho, const view& advection_urho,
                                   const view& advection_vrho, const view& advection_Erho)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    const Span sp_f = sp.merged(sp.shifted(s, s));
    check_view(ustar, "ustar", sp_f);
    check_view(rho, "rho", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(Emat, "Emat", sp);
    check_view(advection_rho, "advection_rho", sp_f, false);
    check_view(advection_urho, "advection_urho", sp_f, false);
    check_view(advection_vrho, "advection_vrho", sp_f, false);
    check_view(advection_Erho, "advection_Erho", sp_f, false);
    if (empty_2d(it)) return;
    // the division by dx: a multiply when dx is a power of two, else the reciprocal-corrected
    // quotient (device_math.hpp), both equal to the IEEE quotient
    int e = 0;
    const bool pow2 = dx > 0 && std::isfinite(dx) && std::frexp(dx, &e) == 0.5;
    const bool recip = !pow2 && dx >= 0x1p-100 && dx <= 0x1p100;
    const flt_t inv_dx = (pow2 || recip) ? flt_t(1) / dx : flt_t(0);
    auto launch = [&](auto kernel) {
        kernel<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, dx, inv_dx, dt, ustar.data, rho.data, umat.data, vmat.data,
                                                     Emat.data, advection_rho.data, advection_urho.data,
                                                     advection_vrho.data, advection_Erho.data);
    };
    if (pow2) launch(k_euler_projection<DX_POW2>);
    else if (recip) launch(k_euler_projection<DX_RECIP>);
    else launch(k_euler_projection<DX_DIVIDE>);
    after_launch("euler_projection");
}
ARMON_KERNEL_CATCH("euler_projection")

ARMON_EXPORT void advection_first_order(const Range& range, const InnerRange2D& inner_range, Idx s, flt_t dt,
                                        const view& ustar, const view& rho, const view& umat, const view& vmat,
                                        const view& Emat, view& advection_rho, view& advection_urho,
                                        view& advection_vrho, view& advection_Erho)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    const Span sp_d = sp.merged(sp.shifted(-s, -s));
    check_view(ustar, "ustar", sp);
    check_view(rho, "rho", sp_d);
    check_view(umat, "umat", sp_d);
    check_view(vmat, "vmat", sp_d);
    check_view(Emat, "Emat", sp_d);
    check_view(advection_rho, "advection_rho", sp, false);
    check_view(advection_urho, "advection_urho", sp, false);
    check_view(advection_vrho, "advection_vrho", sp, false);
    check_view(advection_Erho, "advection_Erho", sp, false);
    if (empty_2d(it)) return;
    k_advection_first_order<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, dt, ustar.data, rho.data, umat.data,
                                                                   vmat.data, Emat.data, advection_rho.data,
                                                                   advection_urho.data, advection_vrho.data,
                                                                   advection_Erho.data);
    after_launch("advection_first_order");
}
ARMON_KERNEL_CATCH("advection_first_order")

ARMON_EXPORT void advection_second_order(const Range& range, const InnerRange2D& inner_range, Idx s, flt_t dx,
                                         flt_t dt, const view& ustar, const view& rho, const view& umat,
                                         const view& vmat, const view& Emat, view& advection_rho,
                                         view& advection_urho, view& advection_vrho, view& advection_Erho)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    // donor d in {i-s, i}; reads ustar at d-s..d+2s and the state at d-s..d+s
    const Span sp_u = sp.shifted(-2 * s, -2 * s).merged(sp.shifted(2 * s, 2 * s)).merged(sp);
    const Span sp_c = sp.shifted(-2 * s, -2 * s).merged(sp.shifted(s, s)).merged(sp);
    check_view(ustar, "ustar", sp_u);
    check_view(rho, "rho", sp_c);
    check_view(umat, "umat", sp_c);
    check_view(vmat, "vmat", sp_c);
    check_view(Emat, "Emat", sp_c);
    check_view(advection_rho, "advection_rho", sp, false);
    check_view(advection_urho, "advection_urho", sp, false);
    check_view(advection_vrho, "advection_vrho", sp, false);
    check_view(advection_Erho, "advection_Erho", sp, false);
    if (empty_2d(it)) return;
    const int64_t rows = it.last_row - it.first_row + 1;
    if (s == 1) {
        // one block per kAdvOut faces of a row
        const dim3 g((unsigned) ((it.len + kAdvOut - 1) / kAdvOut), (unsigned) (rows < 65535 ? rows : 65535));
        k_advection_second_order_x<<<g, kAdvTile, 0, stream()>>>(it, dx, dt, ustar.data, rho.data, umat.data,
                                                                 vmat.data, Emat.data, advection_rho.data,
                                                                 advection_urho.data, advection_vrho.data,
                                                                 advection_Erho.data);
    } else if (s == it.step) {
        // runs of 8 rows (4 while that leaves fewer than 8192 waves of 64 columns): these
        // kernels wait on memory, and short runs keep more waves in flight; the rows a run
        // re-reads before its first output come from L2 (measured at 8192^2: 8 rows 12-16%
        // faster than 64, 4 rows the same as 8)
        int64_t run = 8;
        const int64_t groups = (it.len + 63) / 64;
        if (groups * ((rows + run - 1) / run) < 8192) run = 4;
        const int64_t nruns = (rows + run - 1) / run;
        const dim3 g((unsigned) ((it.len + kBlock - 1) / kBlock), (unsigned) (nruns < 65535 ? nruns : 65535));
        k_advection_second_order_y<<<g, kBlock, 0, stream()>>>(it, run, dx, dt, ustar.data, rho.data, umat.data,
                                                               vmat.data, Emat.data, advection_rho.data,
                                                               advection_urho.data, advection_vrho.data,
                                                               advection_Erho.data);
    } else {
        k_advection_second_order<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, s, dx, dt, ustar.data, rho.data,
                                                                        umat.data, vmat.data, Emat.data,
                                                                        advection_rho.data, advection_urho.data,
                                                                        advection_vrho.data, advection_Erho.data);
    }
    after_launch("advection_second_order");
}
ARMON_KERNEL_CATCH("advection_second_order")

ARMON_EXPORT void boundary_conditions(const Range& range, const InnerRange1D& inner_range, Idx stencil_width,
                                      Idx disp, flt_t u_factor, flt_t v_factor, view& rho, view& umat, view& vmat,
                                      view& pmat, view& cmat, view& gmat, view& Emat)
ARMON_KERNEL_TRY
{
    const Iter1D it = make_iter1d(range, inner_range);
    const Span sp = span_of(it);
    Span all;
    if (stencil_width > 0 && !sp.empty()) {
        // writes i - w*disp, reads i + (w+1)*disp for w in [0, stencil_width)
        const int64_t wl = (stencil_width - 1) * disp;
        all = sp.merged(sp.shifted(-wl, -wl)).merged(sp.shifted(disp, disp)).merged(sp.shifted(disp + wl, disp + wl));
    }
    check_view(rho, "rho", all);
    check_view(umat, "umat", all);
    check_view(vmat, "vmat", all);
    check_view(pmat, "pmat", all);
    check_view(cmat, "cmat", all);
    check_view(gmat, "gmat", all);
    check_view(Emat, "Emat", all);
    if (empty_1d(it) || stencil_width <= 0) return;
    k_boundary_conditions<<<grid_1d(it), kBlock, 0, stream()>>>(it, stencil_width, disp, u_factor, v_factor,
                                                                 rho.data, umat.data, vmat.data, pmat.data, cmat.data,
                                                                 gmat.data, Emat.data);
    after_launch("boundary_conditions");
}
ARMON_KERNEL_CATCH("boundary_conditions")

static Span border_values_span(const Iter2D& it, int64_t side_length, int64_t nghost)
{
    Span s;
    if (it.end <= it.start) return s;
    if (nghost <= 0) throw invalid_argument("nghost must be positive");
    // a(k) = ((k % nghost) * side_length + k / nghost) * 7 + f, f in [0, 7)
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    const int64_t n = it.end - it.start;
    const int64_t probes = n < 2 * nghost + 2 ? n : 2 * nghost + 2;
    auto probe = [&](int64_t k) {
        int64_t a = ((k % nghost) * side_length + k / nghost) * 7;
        lo = a < lo ? a : lo;
        hi = a + 6 > hi ? a + 6 : hi;
    };
    for (int64_t q = 0; q < probes; q++) {
        probe(it.start + q);
        probe(it.end - 1 - q);
    }
    s.add(lo);
    s.add(hi);
    return s;
}

ARMON_EXPORT void read_border_array(const Range& range, const InnerRange2D& inner_range, Idx side_length,
                                    Idx nghost, const view& rho, const view& umat, const view& vmat,
                                    const view& pmat, const view& cmat, const view& gmat, const view& Emat,
                                    view& value_array)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    check_view(rho, "rho", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(pmat, "pmat", sp);
    check_view(cmat, "cmat", sp);
    check_view(gmat, "gmat", sp);
    check_view(Emat, "Emat", sp);
    check_view(value_array, "value_array", border_values_span(it, side_length, nghost), false);
    if (empty_2d(it)) return;
    k_read_border<<<grid_2d(it), kBlock, 0, stream()>>>(it, side_length, nghost, rho.data, umat.data, vmat.data,
                                                         pmat.data, cmat.data, gmat.data, Emat.data,
                                                         value_array.data);
    after_launch("read_border_array");
}
ARMON_KERNEL_CATCH("read_border_array")

ARMON_EXPORT void write_border_array(const Range& range, const InnerRange2D& inner_range, Idx side_length,
                                     Idx nghost, view& rho, view& umat, view& vmat, view& pmat, view& cmat,
                                     view& gmat, view& Emat, const view& value_array)
ARMON_KERNEL_TRY
{
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    check_view(rho, "rho", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(pmat, "pmat", sp);
    check_view(cmat, "cmat", sp);
    check_view(gmat, "gmat", sp);
    check_view(Emat, "Emat", sp);
    check_view(value_array, "value_array", border_values_span(it, side_length, nghost), false);
    if (empty_2d(it)) return;
    k_write_border<<<grid_2d(it), kBlock, 0, stream()>>>(it, side_length, nghost, rho.data, umat.data, vmat.data,
                                                          pmat.data, cmat.data, gmat.data, Emat.data,
                                                          value_array.data);
    after_launch("write_border_array");
}
ARMON_KERNEL_CATCH("write_border_array")

static void launch_init_test(const char* name, const Range& range, const InnerRange2D& inner_range, Idx row_length,
                             Idx nb_ghosts, Idx nx, Idx ny, flt_t sx, flt_t sy, flt_t ox, flt_t oy, Idx pos_x,
                             Idx pos_y, Idx g_nx, Idx g_ny, view& x, view& y, view& rho, view& Emat, view& umat,
                             view& vmat, mask_view& domain_mask, view& pmat, view& cmat, view& gmat, view& ustar,
                             view& pstar, int t, bool debug_indexes, flt_t test_option, const TestState& st)
{
    if (t < 0 || t > 4) {
        char buf[160];
        snprintf(buf, sizeof buf, "Invalid test case index: %d, expected a value between 0 and 4", t);
        throw out_of_range(buf);
    }
    if (row_length <= 0 || nx <= 0 || ny <= 0 || g_nx <= 0 || g_ny <= 0)
        throw invalid_argument("row_length, nx, ny, g_nx and g_ny must be positive");
    const Iter2D it = make_iter2d(range, inner_range);
    const Span sp = span_of(it);
    check_view(x, "x", sp);
    check_view(y, "y", sp);
    check_view(rho, "rho", sp);
    check_view(Emat, "Emat", sp);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(domain_mask, "domain_mask", sp);
    check_view(pmat, "pmat", sp);
    check_view(cmat, "cmat", sp);
    check_view(gmat, "gmat", sp);
    check_view(ustar, "ustar", sp);
    check_view(pstar, "pstar", sp);
    if (empty_2d(it)) return;
    InitArgs a{};
    a.row_length = row_length;
    a.nb_ghosts = nb_ghosts;
    a.nx = nx;
    a.ny = ny;
    a.pos_x = pos_x;
    a.pos_y = pos_y;
    a.g_nx = g_nx;
    a.g_ny = g_ny;
    a.sx = sx;
    a.sy = sy;
    a.ox = ox;
    a.oy = oy;
    a.r = test_option;
    a.test = t;
    a.debug_indexes = debug_indexes;
    a.st = st;
    k_init_test<<<grid_2d_aligned(it), kBlock, 0, stream()>>>(it, a, x.data, y.data, rho.data, Emat.data, umat.data,
                                                       vmat.data, domain_mask.data, pmat.data, cmat.data, gmat.data,
                                                       ustar.data, pstar.data);
    after_launch(name);
}

ARMON_EXPORT void init_test(const Range& range, const InnerRange2D& inner_range, Idx row_length, Idx nb_ghosts,
                            Idx nx, Idx ny, flt_t sx, flt_t sy, flt_t ox, flt_t oy, Idx pos_x, Idx pos_y, Idx g_nx,
                            Idx g_ny, view& x, view& y, view& rho, view& Emat, view& umat, view& vmat,
                            mask_view& domain_mask, view& pmat, view& cmat, view& gmat, view& ustar, view& pstar,
                            Test test, bool debug_indexes, flt_t test_option)
ARMON_KERNEL_TRY
{
    const int t = static_cast<int>(test);
    launch_init_test("init_test", range, inner_range, row_length, nb_ghosts, nx, ny, sx, sy, ox, oy, pos_x, pos_y,
                     g_nx, g_ny, x, y, rho, Emat, umat, vmat, domain_mask, pmat, cmat, gmat, ustar, pstar, t,
                     debug_indexes, test_option, test_state(t, test_option));
}
ARMON_KERNEL_CATCH("init_test")

ARMON_EXPORT void armon_reference_test_state(int test, double test_option, ArmonTestState* out)
{
    const TestState t = test_state(test, (flt_t) test_option);
    *out = ArmonTestState{t.high_rho, t.low_rho, t.high_E, t.low_E, t.high_u, t.low_u, t.high_v, t.low_v};
}

ARMON_EXPORT void init_test_states(const Range& range, const InnerRange2D& inner_range, Idx row_length,
                                   Idx nb_ghosts, Idx nx, Idx ny, flt_t sx, flt_t sy, flt_t ox, flt_t oy, Idx pos_x,
                                   Idx pos_y, Idx g_nx, Idx g_ny, view& x, view& y, view& rho, view& Emat,
                                   view& umat, view& vmat, mask_view& domain_mask, view& pmat, view& cmat,
                                   view& gmat, view& ustar, view& pstar, Test test, flt_t test_option,
                                   const ArmonTestState& states)
ARMON_KERNEL_TRY
{
    TestState st{};
    st.high_rho = (flt_t) states.high_rho;
    st.low_rho = (flt_t) states.low_rho;
    st.high_E = (flt_t) states.high_E;
    st.low_E = (flt_t) states.low_E;
    st.high_u = (flt_t) states.high_u;
    st.low_u = (flt_t) states.low_u;
    st.high_v = (flt_t) states.high_v;
    st.low_v = (flt_t) states.low_v;
    launch_init_test("init_test_states", range, inner_range, row_length, nb_ghosts, nx, ny, sx, sy, ox, oy, pos_x,
                     pos_y, g_nx, g_ny, x, y, rho, Emat, umat, vmat, domain_mask, pmat, cmat, gmat, ustar, pstar,
                     static_cast<int>(test), false, test_option, st);
}
ARMON_KERNEL_CATCH("init_test_states")

static void launch_dt_CFL(const Range& range, const InnerRange1D& inner_range, flt_t dx, flt_t dy,
                          const view& umat, const view& vmat, const view& cmat, const mask_view& domain_mask,
                          double* dt_device)
{
    const Iter1D it = make_iter1d(range, inner_range);
    const Span sp = span_of(it);
    check_view(umat, "umat", sp);
    check_view(vmat, "vmat", sp);
    check_view(cmat, "cmat", sp);
    check_view(domain_mask, "domain_mask", sp);
    Scratch& sc = armon::scratch();
    const double init = (sizeof(flt_t) == 8) ? DBL_MAX : (double) FLT_MAX;
    k_fill_bits<<<1, 1, 0, stream()>>>(sc.min_bits, (unsigned long long) __builtin_bit_cast(long long, init));
    after_launch("dt_CFL(init)");
    if (!empty_1d(it)) {
        int64_t n = it.end - it.start;
        int64_t g = (n + kBlock - 1) / kBlock;
        if (g > kRedBlocks) g = kRedBlocks;
        k_dt_CFL<<<(unsigned) g, kBlock, 0, stream()>>>(it, dx, dy, umat.data, vmat.data, cmat.data, domain_mask.data,
                                                         sc.min_bits);
        after_launch("dt_CFL");
    }
    k_bits_to_double<<<1, 1, 0, stream()>>>(sc.min_bits, dt_device);
    after_launch("dt_CFL(result)");
}

ARMON_EXPORT flt_t dt_CFL(const Range& range, const InnerRange1D& inner_range, flt_t dx, flt_t dy,
                          const view& umat, const view& vmat, const view& cmat, const mask_view& domain_mask)
{
    try {
        Scratch& sc = armon::scratch();
        launch_dt_CFL(range, inner_range, dx, dy, umat, vmat, cmat, domain_mask, sc.result);
        double out = 0;
        hip_check(hipMemcpyAsync(&out, sc.result, sizeof(double), hipMemcpyDeviceToHost, stream()), "dt_CFL copy");
        hip_check(hipStreamSynchronize(stream()), "dt_CFL sync");
        return (flt_t) out;
    }
    catch (const armon::Error& err) { armon::report("dt_CFL", err); }
    catch (const std::exception& err) { armon::report_std("dt_CFL", err); }
    return flt_t(NAN);  // reached only when the handler returns
}

ARMON_EXPORT void dt_CFL_device(const Range& range, const InnerRange1D& inner_range, flt_t dx, flt_t dy,
                                const view& umat, const view& vmat, const view& cmat, const mask_view& domain_mask,
                                double* dt_device)
ARMON_KERNEL_TRY
{
    if (!dt_device) throw invalid_argument("dt_device is null");
    launch_dt_CFL(range, inner_range, dx, dy, umat, vmat, cmat, domain_mask, dt_device);
}
ARMON_KERNEL_CATCH("dt_CFL_device")

// The compensated sums over `range` with `max_blocks` blocks of partials at `partials` and the
// two results at `result` (device memory), returned to the host.
static void run_conservation(const char* name, const Range& range, const InnerRange1D& inner_range, flt_t dx,
                             const view& rho, const view& Emat, const mask_view& domain_mask, double* partials,
                             int max_blocks, double* result, flt_t& total_mass, flt_t& total_energy)
{
    const Iter1D it = make_iter1d(range, inner_range);
    const Span sp = span_of(it);
    check_view(rho, "rho", sp);
    check_view(Emat, "Emat", sp);
    check_view(domain_mask, "domain_mask", sp);
    int64_t n = it.end - it.start;
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > kRedBlocks) g = kRedBlocks;
    if (g < 1) g = 1;
    if (g > max_blocks)
        throw out_of_range(std::string(name) + ": " + std::to_string(g) + " reduction blocks exceed the scratch's "
                           + std::to_string(max_blocks));
    const double ds = (double) dx * (double) dx;
    k_conservation_partial<<<(unsigned) g, kBlock, 0, stream()>>>(it, ds, rho.data, Emat.data, domain_mask.data,
                                                                   partials);
    after_launch(name);
    k_conservation_final<<<1, 1, 0, stream()>>>((int) g, partials, result);
    after_launch(name);
    double out[2];
    hip_check(hipMemcpyAsync(out, result, sizeof out, hipMemcpyDeviceToHost, stream()), "conservation copy");
    hip_check(hipStreamSynchronize(stream()), "conservation sync");
    total_mass = (flt_t) out[0];
    total_energy = (flt_t) out[1];
}

ARMON_EXPORT void conservation_vars(const Range& range, const InnerRange1D& inner_range, flt_t dx, const view& rho,
                                    const view& Emat, const mask_view& domain_mask, flt_t& total_mass,
                                    flt_t& total_energy)
ARMON_KERNEL_TRY
{
    Scratch& sc = armon::scratch();
    run_conservation("conservation_vars", range, inner_range, dx, rho, Emat, domain_mask, sc.partials,
                     Scratch::kPartials / 4, sc.result, total_mass, total_energy);
}
ARMON_KERNEL_CATCH("conservation_vars")

// conservation_vars with its partials and results in the library's static device scratch
// instead of memory allocated on first use: no allocation at any call. The partials are
// written before they are read in every call, so nothing needs clearing; calls that overlap
// on one device (other host threads or streams) share the scratch and race.
ARMON_EXPORT void conservation_vars_noalloc(const Range& range, const InnerRange1D& inner_range, flt_t dx,
                                            const view& rho, const view& Emat, const mask_view& domain_mask,
                                            flt_t& total_mass, flt_t& total_energy)
ARMON_KERNEL_TRY
{
    double* base = noalloc_scratch();
    run_conservation("conservation_vars_noalloc", range, inner_range, dx, rho, Emat, domain_mask, base,
                     kNoAllocBlocks, base + 4 * kNoAllocBlocks, total_mass, total_energy);
}
ARMON_KERNEL_CATCH("conservation_vars_noalloc")

// ---- measurement -----------------------------------------------------------------------

}  // extern "C"

namespace {
// Streaming copy: the achievable HBM bandwidth of this device, against which the sweeps'
// traffic and rate are compared. One 16-byte word per thread and one block per 256 words:
// of the copy shapes measured on MI355X (grid-stride loops of 1-16 Ki blocks with 1-4 words
// in flight per thread, non-temporal or not, and one 8-, 16- or 32-byte word per thread), the
// one-word-per-thread copies of 8 and 16 bytes are the fastest, 6.1-6.3 TB/s.
typedef double copy_word __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) k_copy(const copy_word* __restrict__ in, copy_word* __restrict__ out, int64_t n)
{
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[k];
}
}  // namespace

extern "C" {

ARMON_EXPORT double armon_copy_bandwidth(size_t bytes, int iters)
{
    void* a = nullptr;
    void* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double gbps = -1.0;
    try {
        bytes = bytes / 16 * 16;
        hip_check(hipMalloc(&a, bytes), "hipMalloc(copy)");
        hip_check(hipMalloc(&b, bytes), "hipMalloc(copy)");
        hip_check(hipMemsetAsync(a, 0, bytes, stream()), "hipMemset(copy)");
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const int64_t n = (int64_t) (bytes / 16);
        if ((n + 255) / 256 > INT32_MAX) throw invalid_argument("copy buffer too large");
        const unsigned grid = (unsigned) ((n + 255) / 256);
        k_copy<<<grid, 256, 0, stream()>>>(static_cast<const copy_word*>(a), static_cast<copy_word*>(b), n);
        hip_check(hipEventRecord(e0, stream()), "hipEventRecord");
        for (int i = 0; i < iters; i++)
            k_copy<<<grid, 256, 0, stream()>>>(static_cast<const copy_word*>(a), static_cast<copy_word*>(b), n);
        hip_check(hipEventRecord(e1, stream()), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        gbps = 2.0 * (double) bytes * iters / (ms * 1e-3) / 1e9;
    } catch (const Error& e) {
        report("armon_copy_bandwidth", e);
    }
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    if (a) (void) hipFree(a);
    if (b) (void) hipFree(b);
    return gbps;
}

// utils.cpp:20-27
ARMON_EXPORT int idx_size(void) { return sizeof(Idx); }
ARMON_EXPORT bool is_uidx_signed(void) { return false; }
ARMON_EXPORT int flt_size(void) { return sizeof(flt_t); }

// ---- plumbing -----------------------------------------------------------------------

ARMON_EXPORT int armon_set_stream(void* s)
{
    g_stream = static_cast<hipStream_t>(s);
    return 0;
}

ARMON_EXPORT void* armon_get_stream(void) { return g_stream; }

ARMON_EXPORT int armon_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

ARMON_EXPORT void* armon_malloc(size_t bytes)
{
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        g_last_error = std::string("armon_malloc: ") + hipGetErrorString(e);
        return nullptr;
    }
    return p;
}

#define ARMON_PLUMB(name, expr)                                                      \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            g_last_error = std::string(name) + ": " + hipGetErrorString(e_);        \
            return 1;                                                                \
        }                                                                            \
        return 0;                                                                    \
    } while (0)

ARMON_EXPORT int armon_free(void* p) { ARMON_PLUMB("armon_free", hipFree(p)); }
ARMON_EXPORT int armon_memcpy_htod(void* dst, const void* src, size_t bytes)
{
    ARMON_PLUMB("armon_memcpy_htod", hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
ARMON_EXPORT int armon_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
    ARMON_PLUMB("armon_memcpy_dtoh", hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
ARMON_EXPORT int armon_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
    ARMON_PLUMB("armon_memcpy_dtod", hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
}
ARMON_EXPORT int armon_synchronize(void) { ARMON_PLUMB("armon_synchronize", hipStreamSynchronize(g_stream)); }
ARMON_EXPORT const char* armon_last_error(void) { return g_last_error.c_str(); }
ARMON_EXPORT void armon_clear_error(void) { g_last_error.clear(); }
ARMON_EXPORT const char* armon_version(void) { return "armon-mi355x 0.1 (gfx950)"; }

}  // extern "C"
