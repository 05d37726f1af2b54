// Fused Lagrange + remap sweep kernels (templates), instantiated by fused.hip for FP64 and
// by fused_fp32.hip for the FP32-mixed form, each with its own floating-point flags.
//
// One kernel per axis sweep performs, for every real cell, the whole per-axis chain of the
// reference (armon_2D.cpp:309-320): EOS -> boundary ghost synthesis -> Godunov / GAD face
// fluxes -> Lagrangian cell update -> first/second-order advection fluxes -> Euler
// projection, reading the state once and writing it once (out of place, ping-pong buffers).
//
//  * X sweep (k_sweep_x): one 256-thread block per tile of a stream of padded rows. Each
//    thread owns one cell of the tile (outputs plus an R-cell halo on each side, R = ghost
//    layers read; the batched sweeps can also give each thread 2 cells); the stage results are
//    exchanged through LDS between barriers. Loads and stores are contiguous.
//  * Y sweep (k_sweep_y): one thread per column, marching down a run of rows. Every
//    stage's stencil neighbours along y live in the thread's own registers (a sliding
//    window), so no LDS and no barriers; consecutive lanes touch consecutive columns.
//  * Both number their blocks XCD by XCD (xcd_block) so that each XCD's L2 serves a
//    contiguous part of the arrays.
//  * The batched ensemble (k_batch_x, k_batch_x1, k_batch_y, k_batch_y1) runs the same tile and
//    run bodies over a work list that spans many domains.
//
// Ghost cells are synthesised in registers: layers within the stencil width mirror the real
// cells with the side's velocity factors (halo_exchange.cpp:9-36); deeper layers keep their
// never-refilled values from the arrays, as in the reference.
//
// The CFL reduction of the next cycle (reductions.cpp:8-24) is folded into the epilogue of
// the last sweep of a cycle: it needs the new velocities and the sound speed of that sweep's
// input state, which is exactly what the reference's dt_CFL reads at the start of the next
// cycle (cmat is last written by the EOS of the last sweep).
#pragma once

#include "fused.hpp"
#include "device_math.hpp"
#include "xcd.hpp"

#include <algorithm>
#include <cfloat>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace armon {
namespace sweep {

constexpr int kTile = 256;    // X sweep: cells per tile (FP64 single-domain sweep, batched default)
constexpr int kYBlock = 256;  // Y sweep: columns per block

// Y march: rows per unrolled loop body, the minimum waves per SIMD the register allocation
// must allow (0: the compiler's choice), and whether the FP64 marches load each row's state
// one row before they use it (1) or within the row (0).
#ifndef ARMON_Y_UNROLL
#define ARMON_Y_UNROLL 4
#endif
#ifndef ARMON_Y_WAVES_PER_EU
#define ARMON_Y_WAVES_PER_EU 3   // the FP64 march's occupancy (168 VGPRs); 0: no hint
#endif
#ifndef ARMON_Y_PREFETCH
#define ARMON_Y_PREFETCH 1
#endif
constexpr int kYUnroll = ARMON_Y_UNROLL;
#if ARMON_Y_WAVES_PER_EU > 0
#define ARMON_Y_OCCUPANCY __attribute__((amdgpu_waves_per_eu(ARMON_Y_WAVES_PER_EU)))
#else
#define ARMON_Y_OCCUPANCY
#endif

template <class T>
struct Cell {
    T rho, un, ut, E, p, c;   // un: velocity along the sweep axis, ut: across it
};

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long) b); }
__device__ __forceinline__ unsigned long long double_to_bits(double d) { return (unsigned long long) __double_as_longlong(d); }

__device__ __forceinline__ bool finite_d(double x)
{
    return (double_to_bits(x) & 0x7FF0000000000000ull) != 0x7FF0000000000000ull;
}

// EOS of one cell (fluid_eos.cpp:9-62). The perfect gas is evaluated in the working
// precision T; the Bizarrium EOS always in FP64, because its `e - eps_k0` and `p - pk0`
// terms cancel to a few digits near the reference density (p and c rounded to T).
template <class T, int EOS>
__device__ __forceinline__ void cell_eos(double gamma, T rho, T u, T v, T E, T& p, T& c)
{
    if constexpr (EOS == 0) {
        T g;
        perfect_gas<T>((T) gamma, rho, E, u, v, p, c, g);
    } else {
        double pd, cd, gd;
        bizarrium((double) rho, (double) E, (double) u, (double) v, pd, cd, gd);
        p = (T) pd;
        c = (T) cd;
    }
}

// Cell `k` (along the sweep axis, `n` real cells) of pencil `q` (row for X, column for Y).
// Real cells are loaded; ghost layers 1..stencil_width mirror real cells: the real cell's
// rho, E, p and c, its velocities times the side's factors; deeper ghost layers are read as
// they are, with p and c from stale_p/stale_c.
template <class T, int AXIS, int EOS>
__device__ __forceinline__ Cell<T> load_cell(const SweepArgs<T>& a, int q, int k, int n)
{
    int kk = k;
    int mirror = 0;   // 0 real, 1 low side, 2 high side
    bool stale = false;
    if (k < 0) {
        if (-k <= a.stencil_width) { kk = -k - 1; mirror = 1; }
        else stale = true;
    } else if (k >= n) {
        if (k - n + 1 <= a.stencil_width) { kk = 2 * n - 1 - k; mirror = 2; }
        else stale = true;
    }
    const int idx = AXIS == 0 ? a.index_start + q * a.row_length + kk : a.index_start + kk * a.row_length + q;
    T rho = a.in_rho[idx], u = a.in_u[idx], v = a.in_v[idx], E = a.in_E[idx];
    Cell<T> cell;
    if (stale) {
        cell.p = a.stale_p[idx];
        cell.c = a.stale_c[idx];
    } else {
        // a mirror ghost takes the p and c of its real cell (halo_exchange.cpp:23-30): the EOS
        // of the unscaled state
        cell_eos<T, EOS>(a.gamma, rho, u, v, E, cell.p, cell.c);
    }
    if (mirror == 1) {
        u = u * (T) a.lo_u;
        v = v * (T) a.lo_v;
    } else if (mirror == 2) {
        u = u * (T) a.hi_u;
        v = v * (T) a.hi_v;
    }
    cell.rho = rho;
    cell.E = E;
    cell.un = AXIS == 0 ? u : v;
    cell.ut = AXIS == 0 ? v : u;
    return cell;
}

// load_cell in two parts, for the Y march, which loads a row's state one row before it makes
// the cell: where cell `k` comes from (cell_source: the index it is read at, the mirror side or
// staleness), then the cell from the loaded state (make_cell). The row-ahead loads make the
// FP64 Y sweep 4-9% faster and the FP32-mixed one 2-3% (its 5 waves per SIMD already cover
// most of the load latency).
struct CellSource {
    int idx;
    int mirror;
    bool stale;
};

// STALE = false: the caller guarantees that the stencil width covers every ghost layer read
// (|k| beyond the domain is at most the read extent), so no layer is stale.
template <class T, int AXIS, bool STALE = true>
__device__ __forceinline__ CellSource cell_source(const SweepArgs<T>& a, int q, int k, int n)
{
    int kk = k;
    CellSource s{0, 0, false};
    if (k < 0) {
        if (!STALE || -k <= a.stencil_width) { kk = -k - 1; s.mirror = 1; }
        else s.stale = true;
    } else if (k >= n) {
        if (!STALE || k - n + 1 <= a.stencil_width) { kk = 2 * n - 1 - k; s.mirror = 2; }
        else s.stale = true;
    }
    s.idx = AXIS == 0 ? a.index_start + q * a.row_length + kk : a.index_start + kk * a.row_length + q;
    return s;
}

// The cell from its loaded state (rho, u, v, E at s.idx): EOS (or the stale p and c), then the
// mirror factors.
template <class T, int AXIS, int EOS, bool STALE = true>
__device__ __forceinline__ Cell<T> make_cell(const SweepArgs<T>& a, const CellSource& s, T rho, T u, T v, T E)
{
    Cell<T> cell;
    if (STALE && s.stale) {
        cell.p = a.stale_p[s.idx];
        cell.c = a.stale_c[s.idx];
    } else {
        cell_eos<T, EOS>(a.gamma, rho, u, v, E, cell.p, cell.c);   // unscaled, as load_cell
    }
    if (s.mirror == 1) {
        u = u * (T) a.lo_u;
        v = v * (T) a.lo_v;
    } else if (s.mirror == 2) {
        u = u * (T) a.hi_u;
        v = v * (T) a.hi_v;
    }
    cell.rho = rho;
    cell.E = E;
    cell.un = AXIS == 0 ? u : v;
    cell.ut = AXIS == 0 ? v : u;
    return cell;
}

template <class T>
__device__ __forceinline__ Cell<T> benign_cell()
{
    Cell<T> c;
    c.rho = T(1); c.un = T(0); c.ut = T(0); c.E = T(1); c.p = T(1); c.c = T(1);
    return c;
}

template <class T, int L>
__device__ __forceinline__ void face_flux_gad(T dt, T dx, const Cell<T>& m, const Cell<T>& c, T gu_im, T gp_im,
                                              T gu_i, T gp_i, T gu_ip, T gp_ip, T& us, T& ps)
{
    gad_correct<L, T>(dt, dx, m.rho, c.rho, m.c, c.c, m.un, c.un, m.p, c.p, gu_im, gp_im, gu_i, gp_i, gu_ip, gp_ip,
                      us, ps);
}

__device__ __forceinline__ double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = kmin(v, __shfl_xor(v, off, 64));
    return v;
}

// Merges a per-lane CFL bound into the device accumulator: one atomic per wave.
__device__ __forceinline__ void reduce_cfl(DeviceState* st, double local)
{
    local = wave_min(local);
    if ((threadIdx.x & 63) == 0) atomicMin(&st->dt_min_bits, double_to_bits(local));
}

// ---------------------------------------------------------------------------------------
// X sweep: LDS-staged row tiles
// ---------------------------------------------------------------------------------------

// LDS of one X tile: the cell state, the final face fluxes, and a 4-array scratch that holds
// p, c and the Godunov fluxes until the face fluxes are done (stages 0-2), then the advection
// fluxes (stages 4-5); a barrier separates the last read of the former from the first write
// of the latter. 10 arrays: 20 KiB in FP64, so that 8 blocks fit in a CU's LDS.
template <class T, int TILE = kTile>
struct SmemX {
    T rho[TILE], un[TILE], ut[TILE], E[TILE], us[TILE], ps[TILE];
    T x[4][TILE];
    __device__ T* p() { return x[0]; }
    __device__ T* c() { return x[1]; }
    __device__ T* gu() { return x[2]; }
    __device__ T* gp() { return x[3]; }
    __device__ T (&f())[4][TILE] { return x; }
};

// One X tile of the domain described by `a`, with the cycle's dt. Called by all TILE / CPT
// threads of a block; each thread computes CPT cells of the tile, at slots t, t + TILE / CPT, ...
// (consecutive lanes on consecutive slots for every load and store).
//
// The rows of the domain are laid end to end as a stream of slots, each row taking
// x_slots = nx + 2R slots (its R ghost cells on each side, then the next row). A tile covers
// TILE consecutive slots, starting at slot `k0` of row `j0`; consecutive tiles overlap by 2R
// slots, and the slots [R, TILE - R) of a tile that hold real cells are its outputs. A cell's
// result depends only on the cells within R of it in its own row, so a tile may span the end
// of one row and the start of the next; no thread is spent on a tile's share of a row end.
template <class T, int FLUX, int PROJ, int EOS, int DXM, int TILE = kTile, int CPT = 1>
__device__ __forceinline__ void sweep_x_tile(const SweepArgs<T>& a, T dt, int j0, int k0, SmemX<T, TILE>& sm)
{
    constexpr int GAD = FLUX > 0 ? 1 : 0;
    constexpr int E2 = PROJ == 2 ? 1 : 0;
    constexpr int R = 2 + GAD + E2;
    constexpr int W = TILE - 2 * R;
    constexpr int NT = TILE / CPT;
    static_assert(TILE % CPT == 0, "a tile is a whole number of cells per thread");
    const T dx = (T) a.dx;

    int slot[CPT], jc[CPT], kc[CPT];
    bool in_range[CPT];
    Cell<T> cell[CPT];
    // Row jc and cell kc (-R <= kc < nx + R) of each slot: n = k0 + t < x_slots + TILE. Rows of
    // at least TILE slots: n < 2 x_slots. Shorter rows: the float quotient is within one of the
    // exact one and a single correction makes it exact. (The test on the row length is the
    // block's, taken once, so that only one of the two ways is executed.)
    if (a.x_slots >= TILE) {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int n = k0 + (int) threadIdx.x + c * NT;
            const int dj = n >= a.x_slots ? 1 : 0;
            jc[c] = j0 + dj;
            kc[c] = n - (dj ? a.x_slots : 0) - R;
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int n = k0 + (int) threadIdx.x + c * NT;
            int dj = (int) ((float) n * a.x_inv_slots);
            int r = n - dj * a.x_slots;
            if (r < 0) { dj--; r += a.x_slots; }
            else if (r >= a.x_slots) { dj++; r -= a.x_slots; }
            jc[c] = j0 + dj;
            kc[c] = r - R;
        }
    }
    // Stage 0: load (or synthesise) the cells, EOS.
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int t = threadIdx.x + c * NT;
        slot[c] = t;
        in_range[c] = jc[c] < a.ny;       // false for slots past the last row
        cell[c] = in_range[c] ? load_cell<T, 0, EOS>(a, jc[c], kc[c], a.nx) : benign_cell<T>();
        sm.rho[t] = cell[c].rho;
        sm.un[t] = cell[c].un;
        sm.p()[t] = cell[c].p;
        sm.c()[t] = cell[c].c;
    }
    __syncthreads();

    // Stage 1-2: face flux at face k (between cells k-1 and k).
    T us[CPT], ps[CPT];
    {
        T gu[CPT], gp[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int t = slot[c], tm = t > 0 ? t - 1 : 0;
            godunov<T>(cell[c].rho, sm.rho[tm], cell[c].c, sm.c()[tm], cell[c].un, sm.un[tm], cell[c].p, sm.p()[tm],
                       gu[c], gp[c]);
        }
        if constexpr (GAD) {
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                sm.gu()[slot[c]] = gu[c];
                sm.gp()[slot[c]] = gp[c];
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                const int t = slot[c], tm = t > 0 ? t - 1 : 0, tp = t < TILE - 1 ? t + 1 : TILE - 1;
                Cell<T> m;
                m.rho = sm.rho[tm]; m.un = sm.un[tm]; m.p = sm.p()[tm]; m.c = sm.c()[tm];
                face_flux_gad<T, FLUX - 1>(dt, dx, m, cell[c], sm.gu()[tm], sm.gp()[tm], gu[c], gp[c], sm.gu()[tp],
                                           sm.gp()[tp], us[c], ps[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                us[c] = gu[c];
                ps[c] = gp[c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        sm.us[slot[c]] = us[c];
        sm.ps[slot[c]] = ps[c];
    }
    __syncthreads();

    // Stage 3: Lagrangian update of cell k from faces k and k+1.
    T rho[CPT], un[CPT], ut[CPT], E[CPT], us_p[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int t = slot[c], tp = t < TILE - 1 ? t + 1 : TILE - 1;
        rho[c] = cell[c].rho; un[c] = cell[c].un; ut[c] = cell[c].ut; E[c] = cell[c].E;
        us_p[c] = sm.us[tp];
        cell_update<T>(dx, dt, us[c], us_p[c], ps[c], sm.ps[tp], rho[c], un[c], E[c]);
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int t = slot[c];
        sm.rho[t] = rho[c];
        sm.un[t] = un[c];
        sm.ut[t] = ut[c];
        sm.E[t] = E[c];
    }
    __syncthreads();

    // Stage 4: advection flux at face k.
    T f[CPT][4];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int t = slot[c], tm = t > 0 ? t - 1 : 0, tp = t < TILE - 1 ? t + 1 : TILE - 1;
        const T disp = dt * us[c];
        if constexpr (E2) {
            T Dx;
            int td;
            if (disp > 0) {
                Dx = -(dx - dt * sm.us[tm]);
                td = tm;
            } else {
                Dx = dx + dt * sm.us[tp];
                td = t;
            }
            const int tdm = td > 0 ? td - 1 : 0;
            const int tdp = td < TILE - 1 ? td + 1 : TILE - 1;
            const int tdpp = td < TILE - 2 ? td + 2 : TILE - 1;
            advection_flux_2<T>(dx, dt, disp, Dx, sm.us[tdm], sm.us[td], sm.us[tdp], sm.us[tdpp], sm.rho[tdm],
                                sm.rho[td], sm.rho[tdp], sm.un[tdm], sm.un[td], sm.un[tdp], sm.ut[tdm], sm.ut[td],
                                sm.ut[tdp], sm.E[tdm], sm.E[td], sm.E[tdp], f[c][0], f[c][1], f[c][2], f[c][3]);
        } else {
            const int td = disp > 0 ? tm : t;
            advection_flux_1<T>(disp, sm.rho[td], sm.un[td], sm.ut[td], sm.E[td], f[c][0], f[c][1], f[c][2], f[c][3]);
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
#pragma unroll
        for (int q = 0; q < 4; q++) sm.f()[q][slot[c]] = f[c][q];
    }
    __syncthreads();

    // Stage 5: projection of cell k, output.
    double local = DBL_MAX;
    bool any = false;
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int t = slot[c], tp = t < TILE - 1 ? t + 1 : TILE - 1;
        if (t >= R && t < R + W && in_range[c] && kc[c] >= 0 && kc[c] < a.nx) {
            projection<DXM, T>(dx, (T) a.inv_dx, dt, us[c], us_p[c], f[c][0], sm.f()[0][tp], f[c][1], sm.f()[1][tp],
                               f[c][2], sm.f()[2][tp], f[c][3], sm.f()[3][tp], rho[c], un[c], ut[c], E[c]);
            const int idx = a.index_start + jc[c] * a.row_length + kc[c];
            a.out_rho[idx] = rho[c];
            a.out_u[idx] = un[c];
            a.out_v[idx] = ut[c];
            a.out_E[idx] = E[c];
            if (a.reduce) {
                const double v = cfl_cell<double>(a.cfl_dx, a.cfl_dy, (double) un[c], (double) ut[c],
                                                  (double) cell[c].c, 1.0);
                local = any ? kmin(local, v) : v;
                any = true;
            }
        }
    }
    if (a.reduce) reduce_cfl(a.st, local);
}

// Tile of the single-domain X sweep per element type: FP64 tiles of 256 cells at one cell per
// thread; FP32 tiles of 512 cells at two cells per thread (256 threads, the same 20 KiB of LDS),
// which measured faster for the FP32-mixed sweep and slower for the FP64 one, whose 3 VALU
// instructions of 4 are FP64 arithmetic with little latency left to hide.
template <class T>
struct XTile {
    static constexpr int cells = 256, cpt = 1;
};
template <>
struct XTile<float> {
    static constexpr int cells = 512, cpt = 2;
};

template <class T, int FLUX, int PROJ, int EOS, int DXM>
__global__ void __launch_bounds__(XTile<T>::cells / XTile<T>::cpt) k_sweep_x(SweepArgs<T> a)
{
    constexpr int TILE = XTile<T>::cells;
    constexpr int W = TILE - 2 * (2 + (FLUX > 0 ? 1 : 0) + (PROJ == 2 ? 1 : 0));
    if (!a.st->active) return;
    const T dt = (T) (a.st->current_dt * a.dt_factor);
    __shared__ SmemX<T, TILE> sm;
    const int s0 = xcd_block(blockIdx.x, gridDim.x) * W;   // first slot of the tile
    const int j0 = s0 / a.x_slots;
    sweep_x_tile<T, FLUX, PROJ, EOS, DXM, TILE, XTile<T>::cpt>(a, dt, j0, s0 - j0 * a.x_slots, sm);
}

// ---------------------------------------------------------------------------------------
// Y sweep: register march down columns
// ---------------------------------------------------------------------------------------

template <class T, int N>
__device__ __forceinline__ void push(T (&w)[N], T v)
{
#pragma unroll
    for (int i = 0; i < N - 1; i++) w[i] = w[i + 1];
    w[N - 1] = v;
}

// The march state of one column of a Y run: sliding windows of the stage results along y.
// AHEAD: each row's state is loaded one row before it is used (the FP64 sweeps); with
// STALE = false, when the stencil width covers the read extent so that no ghost layer is stale,
// the march then loads nothing else.
// Whether the Y march shares divisors between Godunov's two quotients and the projection's three
// (with_divisor), as the X tiles do (3 of their 17 FP64 reciprocals per cell removed, 3% of
// their time). While the march's waves ran independently it gained nothing (4% fewer VALU busy
// cycles, the same time: the waves waited on their dependency chains); with the block barrier
// every 4 rows it is 2.3% faster at 8192^2 and still fits 168 VGPRs with no spill.
constexpr bool kYSharedDivisor = true;

template <class T, int FLUX, int PROJ, int EOS, int DXM, bool STALE, bool AHEAD>
struct YColumn {
    static constexpr int GAD = FLUX > 0 ? 1 : 0;
    static constexpr int E2 = PROJ == 2 ? 1 : 0;
    static constexpr int R = 2 + GAD + E2;

    // Windows, oldest first (see the index map in each stage of step()).
    Cell<T> pre[3];                       // cells n-2, n-1, n (pre-update state + EOS)
    T g_u[3], g_p[3];                     // Godunov at faces n-2, n-1, n
    T f_u[5], f_p[2];                     // final face flux: u* at faces F-4..F, p* at F-1..F (F = n - GAD)
    T q_rho[4], q_un[4], q_ut[4], q_E[4]; // post-update cells U-3..U (U = F - 1)
    T h0[2], h1[2], h2[2], h3[2];         // advection fluxes at faces A-1, A (A = U - E2)
    T nx_rho, nx_u, nx_v, nx_E;           // the next row's state (AHEAD)
    int x;
    bool lane_ok;                         // false for lanes past the domain
    double local;                         // CFL bound of the column's outputs

    __device__ __forceinline__ void init(int x_, bool ok)
    {
        x = x_;
        lane_ok = ok;
#pragma unroll
        for (int i = 0; i < 3; i++) { pre[i] = benign_cell<T>(); g_u[i] = T(0); g_p[i] = T(0); }
#pragma unroll
        for (int i = 0; i < 5; i++) f_u[i] = T(0);
#pragma unroll
        for (int i = 0; i < 4; i++) { q_rho[i] = T(1); q_un[i] = T(0); q_ut[i] = T(0); q_E[i] = T(1); }
        f_p[0] = f_p[1] = T(0);
        h0[0] = h0[1] = h1[0] = h1[1] = h2[0] = h2[1] = h3[0] = h3[1] = T(0);
        nx_rho = T(1); nx_u = T(0); nx_v = T(0); nx_E = T(1);
        local = DBL_MAX;
    }

    __device__ __forceinline__ void fetch(const SweepArgs<T>& a, int k)
    {
        if (lane_ok) {
            const int idx = cell_source<T, 1, STALE>(a, x, k, a.ny).idx;
            nx_rho = a.in_rho[idx];
            nx_u = a.in_u[idx];
            nx_v = a.in_v[idx];
            nx_E = a.in_E[idx];
        }
    }

    // Row n of a run that outputs rows [y0, y1). `ring` (LDS, 8 rows x the block's columns) keeps
    // each row's sound speed until the CFL bound of its output, R rows later: off the arithmetic's
    // dependency chains, and 8 VGPRs fewer than a register window.
    __device__ __forceinline__ void step(const SweepArgs<T>& a, T dt, T dx, T inv_dx, int n, int y0, int y1,
                                         T (*ring)[kYBlock])
    {
        Cell<T> cn;
        if constexpr (AHEAD) {
            // the next load is issued after this row's cell is made: a stale layer's p, c loads
            // are waited for there
            cn = lane_ok ? make_cell<T, 1, EOS, STALE>(a, cell_source<T, 1, STALE>(a, x, n, a.ny), nx_rho, nx_u, nx_v,
                                                       nx_E)
                         : benign_cell<T>();
            if (n + 1 < y1 + R) fetch(a, n + 1);
        } else {
            cn = lane_ok ? load_cell<T, 1, EOS>(a, x, n, a.ny) : benign_cell<T>();
        }
        pre[0] = pre[1];
        pre[1] = pre[2];
        pre[2] = cn;
        ring[n & 7][threadIdx.x] = cn.c;

        // Godunov at face n (cells n-1 | n).
        T gu, gp;
        godunov<T, kYSharedDivisor>(pre[2].rho, pre[1].rho, pre[2].c, pre[1].c, pre[2].un, pre[1].un, pre[2].p,
                                    pre[1].p, gu, gp);
        push(g_u, gu);
        push(g_p, gp);

        // Final flux at face F: GAD at F = n-1 (cells n-2 | n-1, Godunov n-2, n-1, n), else F = n.
        T us, ps;
        if constexpr (GAD) {
            face_flux_gad<T, FLUX - 1>(dt, dx, pre[0], pre[1], g_u[0], g_p[0], g_u[1], g_p[1], g_u[2], g_p[2], us,
                                       ps);
        } else {
            us = gu;
            ps = gp;
        }
        push(f_u, us);
        push(f_p, ps);

        // Lagrangian update of cell U = F-1 from faces U (f_*[3], f_p[0]) and F (f_*[4], f_p[1]).
        const Cell<T>& pu = GAD ? pre[0] : pre[1];
        T rho = pu.rho, un = pu.un, ut = pu.ut, E = pu.E;
        cell_update<T>(dx, dt, f_u[3], f_u[4], f_p[0], f_p[1], rho, un, E);
        push(q_rho, rho);
        push(q_un, un);
        push(q_ut, ut);
        push(q_E, E);

        // Advection flux at face A.
        T fa0, fa1, fa2, fa3;
        if constexpr (E2) {
            // A = U-1 = F-2: u*[A] = f_u[2]; cells q[k] = U-3+k.
            const T disp = dt * f_u[2];
            if (disp > 0) {
                const T Dx = -(dx - dt * f_u[1]);
                // donor d = A-1 = q[1]; u* at d-1..d+2 = f_u[0..3]
                advection_flux_2<T>(dx, dt, disp, Dx, f_u[0], f_u[1], f_u[2], f_u[3], q_rho[0], q_rho[1], q_rho[2],
                                    q_un[0], q_un[1], q_un[2], q_ut[0], q_ut[1], q_ut[2], q_E[0], q_E[1], q_E[2], fa0,
                                    fa1, fa2, fa3);
            } else {
                const T Dx = dx + dt * f_u[3];
                // donor d = A = q[2]; u* at d-1..d+2 = f_u[1..4]
                advection_flux_2<T>(dx, dt, disp, Dx, f_u[1], f_u[2], f_u[3], f_u[4], q_rho[1], q_rho[2], q_rho[3],
                                    q_un[1], q_un[2], q_un[3], q_ut[1], q_ut[2], q_ut[3], q_E[1], q_E[2], q_E[3], fa0,
                                    fa1, fa2, fa3);
            }
        } else {
            // A = U = F-1: u*[A] = f_u[3]; donor A-1 = q[2] or A = q[3].
            const T disp = dt * f_u[3];
            if (disp > 0)
                advection_flux_1<T>(disp, q_rho[2], q_un[2], q_ut[2], q_E[2], fa0, fa1, fa2, fa3);
            else
                advection_flux_1<T>(disp, q_rho[3], q_un[3], q_ut[3], q_E[3], fa0, fa1, fa2, fa3);
        }
        push(h0, fa0);
        push(h1, fa1);
        push(h2, fa2);
        push(h3, fa3);

        // Projection of cell P = A-1 = n-R from faces P (h*[0]) and A (h*[1]).
        const int P = n - R;
        if (P >= y0 && P < y1 && lane_ok) {
            constexpr int QP = E2 ? 1 : 2;    // q index of cell P
            constexpr int FP = E2 ? 1 : 2;    // f_u index of face P
            T r = q_rho[QP], u = q_un[QP], v = q_ut[QP], e = q_E[QP];
            projection<DXM, T, kYSharedDivisor>(dx, inv_dx, dt, f_u[FP], f_u[FP + 1], h0[0], h0[1], h1[0], h1[1],
                                                h2[0], h2[1], h3[0], h3[1], r, u, v, e);
            const int idx = a.index_start + P * a.row_length + x;
            a.out_rho[idx] = r;
            a.out_u[idx] = v;   // ut is the x velocity in a Y sweep
            a.out_v[idx] = u;
            a.out_E[idx] = e;
            if (a.reduce)
                local = kmin(local, cfl_cell<double>(a.cfl_dx, a.cfl_dy, (double) v, (double) u,
                                                     (double) ring[P & 7][threadIdx.x], 1.0));
        }
    }
};

// One Y run: column x (a lane at or past nx marches benign cells and writes nothing), output
// rows [y0, y1), with the cycle's dt. The whole wave must call it with the same y0, y1.
//
// SYNC: every wave of the block marches the same rows (of adjacent columns) and meets a block
// barrier after every 4 rows, so that the block's loads and stores of a row stay within 4 rows
// of each other. Waves that march independently drift apart by many rows, and the pieces that
// each wave reads and writes per array and row (512 bytes in FP64) then fall in different DRAM
// pages: a copy with the FP64 march's access pattern and no arithmetic (ymarch_copy.hip) moves
// 4.2 TB/s that way and 4.6 TB/s with a barrier every 1-8 rows. With the barrier the FP32-mixed
// march (98 VGPRs) is 9% faster and the FP64 march 3.6%; the FP64 march fits 3 waves per SIMD
// with it only with the sound-speed window in LDS and the 3-wave occupancy hint (168 VGPRs,
// against 178 with a register window). `idle` waves (past the last column) take part in the
// barriers only. SYNC needs the same number of rows for the whole block: the loop-free batched
// kernel takes it for a block whose 4 waves hold live runs of equal length (most blocks, as
// the runs are listed longest first) and the independent march otherwise.
template <class T, int FLUX, int PROJ, int EOS, int DXM, bool STALE, bool AHEAD, bool SYNC = false>
__device__ __forceinline__ void sweep_y_run(const SweepArgs<T>& a, T dt, int x, int y0, int y1, bool idle = false)
{
    using Col = YColumn<T, FLUX, PROJ, EOS, DXM, STALE, AHEAD>;
    constexpr int R = Col::R;
    // a row counter from 0 keeps the loop and the row index in scalar registers (counting
    // from y0 - R, the compiler carried the row index in a vector register)
    const int rows = y1 - y0 + 2 * R;
    // SYNC marches whole groups of 4 rows: up to 3 steps past the run, which load nothing (the
    // row-ahead load stops at the run's last row; without stale layers a cell is made from the
    // loaded state alone) and store nothing (outputs are rows [y0, y1) only)
    static_assert(!(SYNC && STALE), "the block-synchronised march has no stale ghost layers");
    const int groups = (rows + 3) / 4;
    if constexpr (SYNC) {
        if (idle) {
            for (int g = 0; g < groups; ++g) __syncthreads();
            return;
        }
    }
    Col col;
    __shared__ T ring[8][kYBlock];
    col.init(x, x < a.nx);
    if constexpr (AHEAD) col.fetch(a, y0 - R);
    if constexpr (SYNC) {
        for (int g = 0; g < groups; ++g) {
#pragma unroll
            for (int k = 0; k < 4; ++k) col.step(a, dt, (T) a.dx, (T) a.inv_dx, y0 - R + 4 * g + k, y0, y1, ring);
            __syncthreads();
        }
    } else {
#pragma unroll kYUnroll
        for (int i = 0; i < rows; ++i) col.step(a, dt, (T) a.dx, (T) a.inv_dx, y0 - R + i, y0, y1, ring);
    }
    if (a.reduce) reduce_cfl(a.st, col.local);
}

template <class T, int FLUX, int PROJ, int EOS, int DXM, bool STALE>
__global__ void __launch_bounds__(kYBlock) ARMON_Y_OCCUPANCY k_sweep_y(SweepArgs<T> a, int ncolblk)
{
    if (!a.st->active) return;
    const T dt = (T) (a.st->current_dt * a.dt_factor);
    const int bid = xcd_block(blockIdx.x, gridDim.x);   // runs in segment-major order
    const int seg = bid / ncolblk;
    const int x = (bid - seg * ncolblk) * kYBlock + threadIdx.x;
    const bool idle = x - (int) (threadIdx.x & 63) >= a.nx;   // a wave past the last column
    const int y0 = seg * a.seg_len;
    constexpr bool ahead = ARMON_Y_PREFETCH;
    constexpr bool sync = !STALE;   // see sweep_y_run
    if (idle && !sync) return;   // no barriers follow
    sweep_y_run<T, FLUX, PROJ, EOS, DXM, STALE, ahead, sync>(a, dt, x, y0, min(y0 + a.seg_len, a.ny), idle);
}

// ---------------------------------------------------------------------------------------
// Batched ensemble: persistent blocks over the work items of every domain
// ---------------------------------------------------------------------------------------

// How the batched sweeps divide by the cell sizes, the same for every domain: the
// reciprocal-corrected quotient in FP64 (exact for powers of two too), the FP32-mixed division
// (a native reciprocal) in FP32.
template <class T>
constexpr int kBatchDxMode = sizeof(T) == 8 ? DX_RECIP : DX_DIVIDE;

template <class T>
__device__ __forceinline__ SweepArgs<T> batch_domain_args(const BatchArgs<T>& b, int d, int axis)
{
    const BatchDomain& D = b.domains[d];
    SweepArgs<T> a;
    a.in_rho = b.in_rho + D.offset;
    a.in_u = b.in_u + D.offset;
    a.in_v = b.in_v + D.offset;
    a.in_E = b.in_E + D.offset;
    a.out_rho = b.out_rho + D.offset;
    a.out_u = b.out_u + D.offset;
    a.out_v = b.out_v + D.offset;
    a.out_E = b.out_E + D.offset;
    a.stale_p = b.stale_p + D.offset;
    a.stale_c = b.stale_c + D.offset;
    a.st = b.states + d;
    a.nx = D.nx;
    a.ny = D.ny;
    a.row_length = D.row_length;
    a.index_start = D.index_start;
    a.stencil_width = b.stencil_width;
    a.seg_len = 0;
    a.reduce = b.reduce;
    a.gamma = b.gamma;
    a.dx = D.dx[axis];
    a.inv_dx = D.inv_dx[axis];
    a.dt_factor = b.dt_factor;
    const int lo = axis == 0 ? 0 : 3, hi = axis == 0 ? 1 : 2;   // Left/Right or Bottom/Top
    a.lo_u = D.bc[lo][0];
    a.lo_v = D.bc[lo][1];
    a.hi_u = D.bc[hi][0];
    a.hi_v = D.bc[hi][1];
    a.cfl_dx = D.cfl_dx;
    a.cfl_dy = D.cfl_dy;
    a.x_slots = D.x_slots;
    a.x_inv_slots = D.x_inv_slots;
    return a;
}

template <class T, int FLUX, int PROJ, int EOS, int TILE>
__global__ void __launch_bounds__(TILE) k_batch_x(BatchArgs<T> b, const BatchItemX* __restrict__ items, int nitems)
{
    __shared__ SmemX<T, TILE> sm;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BatchItemX item = items[it];
        const DeviceState* st = b.states + item.domain;
        if (!st->active) continue;   // the same for the whole block
        const SweepArgs<T> a = batch_domain_args(b, item.domain, 0);
        sweep_x_tile<T, FLUX, PROJ, EOS, kBatchDxMode<T>, TILE>(a, (T) (st->current_dt * b.dt_factor), item.row,
                                                               item.slot, sm);
        __syncthreads();             // the next item reuses the LDS
    }
}

// One block per X item (the launch configuration the autotuner usually picks): no item loop,
// so the tile body keeps the single-domain kernel's register budget. TILE cells per tile,
// CPT cells per thread.
template <class T, int FLUX, int PROJ, int EOS, int TILE, int CPT = 1>
__global__ void __launch_bounds__(TILE / CPT) k_batch_x1(BatchArgs<T> b, const BatchItemX* __restrict__ items)
{
    __shared__ SmemX<T, TILE> sm;
    const BatchItemX item = items[xcd_block(blockIdx.x, gridDim.x)];
    const DeviceState* st = b.states + item.domain;
    if (!st->active) return;
    const SweepArgs<T> a = batch_domain_args(b, item.domain, 0);
    sweep_x_tile<T, FLUX, PROJ, EOS, kBatchDxMode<T>, TILE, CPT>(a, (T) (st->current_dt * b.dt_factor), item.row,
                                                                item.slot, sm);
}

// The row-ahead loads of the FP64 march need 179 VGPRs here (the per-item domain arguments
// add to the single-domain sweep's 164), which would leave 2 waves per SIMD; capped at 3 waves,
// the compiler keeps 7 dwords of item-invariant values in scratch instead (one 16-byte reload
// per row), and the FP64 sweep is 4% faster than without the row-ahead loads.
template <class T, int FLUX, int PROJ, int EOS, bool STALE>
__global__ void __launch_bounds__(kYBlock) __attribute__((amdgpu_waves_per_eu(3)))
        k_batch_y(BatchArgs<T> b, const BatchItemY* __restrict__ items, int nitems)
{
    const int lane = threadIdx.x & 63;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    // the item index is uniform across the wave: keep it (and what is loaded through it) scalar
    const int first = __builtin_amdgcn_readfirstlane((int) ((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    for (int it = first; it < nitems; it += nwaves) {
        const BatchItemY item = items[it];
        const DeviceState* st = b.states + item.domain;
        if (!st->active) continue;   // the same for the whole wave
        const SweepArgs<T> a = batch_domain_args(b, item.domain, 1);
        const int x = item.x0 + lane;
        constexpr bool ahead = ARMON_Y_PREFETCH;
        sweep_y_run<T, FLUX, PROJ, EOS, kBatchDxMode<T>, STALE, ahead>(a, (T) (st->current_dt * b.dt_factor), x,
                                                                       item.y0, item.y1);
    }
}

// One wave per Y item (the launch configuration the autotuner usually picks): no item loop, so
// the run body keeps the single-domain kernel's register budget. The work list is sorted by run
// length and otherwise keeps each domain's runs in column order, so a block's 4 waves are mostly
// adjacent columns of one run: when all 4 waves have active items of the same row count, the
// block marches them with the block barrier of sweep_y_run (same trip count for every wave),
// which keeps their accesses of a row together in DRAM.
template <class T, int FLUX, int PROJ, int EOS, bool STALE>
__global__ void __launch_bounds__(kYBlock) __attribute__((amdgpu_waves_per_eu(3)))
        k_batch_y1(BatchArgs<T> b, const BatchItemY* __restrict__ items, int nitems)
{
    const int lane = threadIdx.x & 63;
    const int it = __builtin_amdgcn_readfirstlane((int) ((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    BatchItemY item{0, 0, 0, 0};
    bool live = it < nitems;
    if (live) {
        item = items[it];
        live = b.states[item.domain].active != 0;
    }
    bool sync = false;
    if constexpr (!STALE) {
        __shared__ int rows_of_wave[kYBlock / 64];
        if (lane == 0) rows_of_wave[threadIdx.x >> 6] = live ? item.y1 - item.y0 : -1;
        __syncthreads();
        sync = true;
#pragma unroll
        for (int w = 0; w < kYBlock / 64; w++) sync = sync && rows_of_wave[w] == rows_of_wave[0] && rows_of_wave[w] >= 0;
    }
    if (!live) return;   // the whole wave; never in a synchronised block
    const DeviceState* st = b.states + item.domain;
    const SweepArgs<T> a = batch_domain_args(b, item.domain, 1);
    const int x = item.x0 + lane;
    constexpr bool ahead = ARMON_Y_PREFETCH;
    const T dt = (T) (st->current_dt * b.dt_factor);
    if constexpr (!STALE) {
        if (sync) {
            sweep_y_run<T, FLUX, PROJ, EOS, kBatchDxMode<T>, STALE, ahead, true>(a, dt, x, item.y0, item.y1);
            return;
        }
    }
    sweep_y_run<T, FLUX, PROJ, EOS, kBatchDxMode<T>, STALE, ahead>(a, dt, x, item.y0, item.y1);
}

// X tiles covering a stream of `slots` slots (tiles of W outputs overlapping by 2R): the last
// real cell sits R slots before the end.
inline int64_t x_tiles(int64_t slots, int R, int W) { return (slots - 2 * R + W - 1) / W; }

inline unsigned grid_for(int64_t n, int block, int64_t max_blocks)
{
    int64_t g = (n + block - 1) / block;
    if (g > max_blocks) g = max_blocks;
    if (g < 1) g = 1;
    return (unsigned) g;
}

// Blocks of `block` threads of `kernel` resident on the whole device at once.
inline int resident_blocks(const void* kernel, int block)
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) return 0;
    return per_cu * cus;
}

// SIMDs of the current device (4 per CU), looked up once per device.
inline int device_simds()
{
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    if (cached[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            return 1024;
        cached[dev] = 4 * cus;
    }
    return cached[dev];
}

// Rows per marching run of the Y sweep, for a domain of nx x ny cells whose stencil reads
// `halo` cells each side. A run of L output rows marches L + 2 halo rows (the warm-up), and a
// wave of 64 columns marches one run, so shorter runs give more waves for more redundant work.
// The rate of the march grows with the waves in flight as w / (w + SIMDs) (measured on MI355X
// from 512^2 to 8192^2: the best L is 8 rows at 512^2, 16 at 1024^2, 32 at 2048^2, 32-64 at
// 4096^2 and 64-128 at 8192^2), so L minimises (1 + 2 halo / L) (1 + SIMDs / waves) over the
// powers of two from 8 to 128, the larger L on a tie. The number of runs per column is a
// multiple of the XCD count when the column has at least that many, so that each XCD's
// contiguous share of the blocks (xcd_block) is a set of whole row bands.
inline int y_run_rows(int nx, int ny, int halo)
{
    if (ny <= 8) return ny > 0 ? ny : 1;
    const double simds = (double) device_simds();
    const double groups = (double) ((nx + 63) / 64);
    int best = ny;
    double best_cost = 0;
    for (int L = 8; L <= 128; L *= 2) {
        int nseg = (ny + L - 1) / L;
        if (nseg >= kXcds) nseg = (nseg + kXcds - 1) / kXcds * kXcds;
        const int len = (ny + nseg - 1) / nseg;
        nseg = (ny + len - 1) / len;
        const double cost = (1.0 + 2.0 * halo / len) * (1.0 + simds / (groups * nseg));
        if (L == 8 || cost <= best_cost * (1.0 + 1e-3)) {
            best = len;
            best_cost = cost;
        }
    }
    return best;
}

inline void check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T, int F, int P, int E, int DXM>
void launch_variant(int axis, const SweepArgs<T>& a, hipStream_t s)
{
    constexpr int R = 2 + (F > 0 ? 1 : 0) + (P == 2 ? 1 : 0);
    if (axis == 0) {
        constexpr int W = XTile<T>::cells - 2 * R;
        SweepArgs<T> b = a;
        b.x_slots = a.nx + 2 * R;
        b.x_inv_slots = 1.0f / (float) b.x_slots;
        const int64_t tiles = x_tiles((int64_t) a.ny * b.x_slots, R, W);
        k_sweep_x<T, F, P, E, DXM><<<dim3((unsigned) tiles), dim3(XTile<T>::cells / XTile<T>::cpt), 0, s>>>(b);
        check_launch("sweep_x");
    } else {
        // Each block marches down a run of rows of its 256 columns (the warm-up is paid once
        // per run); blocks are numbered row band by row band (segment-major).
        const int ncolblk = (a.nx + kYBlock - 1) / kYBlock;
        SweepArgs<T> b = a;
        b.seg_len = a.seg_len > 0 ? std::min(a.seg_len, a.ny) : y_run_rows(a.nx, a.ny, R);
        const int nseg = (a.ny + b.seg_len - 1) / b.seg_len;
        const dim3 grid((unsigned) (ncolblk * nseg));
        if (a.stencil_width >= R)
            k_sweep_y<T, F, P, E, DXM, false><<<grid, dim3(kYBlock), 0, s>>>(b, ncolblk);
        else
            k_sweep_y<T, F, P, E, DXM, true><<<grid, dim3(kYBlock), 0, s>>>(b, ncolblk);
        check_launch("sweep_y");
    }
}

// The division by dx of the single-domain sweeps: a multiply for a power of two, else the
// reciprocal-corrected quotient in FP64 and the FP32-mixed division in FP32.
template <class T, int F, int P, int E>
void launch_dx_dispatch(int axis, const SweepArgs<T>& a, hipStream_t s)
{
    constexpr int other = sizeof(T) == 8 ? DX_RECIP : DX_DIVIDE;
    if (a.dx_mode == DX_POW2) launch_variant<T, F, P, E, DX_POW2>(axis, a, s);
    else launch_variant<T, F, P, E, other>(axis, a, s);
}

template <class T, int F, int P>
void launch_eos_dispatch(int axis, int eos, const SweepArgs<T>& a, hipStream_t s)
{
    if (eos == 0) launch_dx_dispatch<T, F, P, 0>(axis, a, s);
    else launch_dx_dispatch<T, F, P, 1>(axis, a, s);
}

template <class T, int F>
void launch_proj_dispatch(int axis, const SweepVariant& v, const SweepArgs<T>& a, hipStream_t s)
{
    if (v.proj == 2) launch_eos_dispatch<T, F, 2>(axis, v.eos, a, s);
    else launch_eos_dispatch<T, F, 1>(axis, v.eos, a, s);
}

template <class T, int F, int P, int E, int TILE>
void launch_batch_x(const BatchArgs<T>& b, const BatchItemX* items, int nitems, int grid, hipStream_t s)
{
    if (grid >= nitems)
        k_batch_x1<T, F, P, E, TILE><<<dim3((unsigned) nitems), dim3(TILE), 0, s>>>(b, items);
    else
        k_batch_x<T, F, P, E, TILE><<<dim3((unsigned) grid), dim3(TILE), 0, s>>>(b, items, nitems);
}

// Two cells per thread: one block per tile of TILE cells (TILE / 2 threads).
template <class T, int F, int P, int E, int TILE>
void launch_batch_x2(const BatchArgs<T>& b, const BatchItemX* items, int nitems, hipStream_t s)
{
    k_batch_x1<T, F, P, E, TILE, 2><<<dim3((unsigned) nitems), dim3(TILE / 2), 0, s>>>(b, items);
}

// X: `tile` cells per tile (128, 256 or 512), `cpt` cells per thread (1; or 2 for tiles of 256
// or 512 cells, one block per tile); `grid` blocks, or one per tile when grid >= nitems.
template <class T, int F, int P, int E>
void launch_batch_variant(int axis, const BatchArgs<T>& b, const void* items, int nitems, int grid, int tile, int cpt,
                          hipStream_t s)
{
    if (nitems <= 0) return;
    if (axis == 0) {
        const auto* x_items = static_cast<const BatchItemX*>(items);
        if (cpt == 2) {
            if (tile == 512) launch_batch_x2<T, F, P, E, 512>(b, x_items, nitems, s);
            else if (tile == 256) launch_batch_x2<T, F, P, E, 256>(b, x_items, nitems, s);
            else throw std::invalid_argument("2 cells per thread needs X tiles of 256 or 512 cells");
        } else {
            switch (tile) {
            case 128: launch_batch_x<T, F, P, E, 128>(b, x_items, nitems, grid, s); break;
            case 512: launch_batch_x<T, F, P, E, 512>(b, x_items, nitems, grid, s); break;
            default: launch_batch_x<T, F, P, E, 256>(b, x_items, nitems, grid, s); break;
            }
        }
    } else {
        constexpr int R = 2 + (F > 0 ? 1 : 0) + (P == 2 ? 1 : 0);
        const auto* y_items = static_cast<const BatchItemY*>(items);
        const int waves_per_block = kYBlock / 64;
        if ((int64_t) grid * waves_per_block >= nitems) {
            const unsigned g1 = (unsigned) ((nitems + waves_per_block - 1) / waves_per_block);
            if (b.stencil_width >= R)
                k_batch_y1<T, F, P, E, false><<<dim3(g1), dim3(kYBlock), 0, s>>>(b, y_items, nitems);
            else
                k_batch_y1<T, F, P, E, true><<<dim3(g1), dim3(kYBlock), 0, s>>>(b, y_items, nitems);
        } else if (b.stencil_width >= R) {
            k_batch_y<T, F, P, E, false><<<dim3((unsigned) grid), dim3(kYBlock), 0, s>>>(b, y_items, nitems);
        } else {
            k_batch_y<T, F, P, E, true><<<dim3((unsigned) grid), dim3(kYBlock), 0, s>>>(b, y_items, nitems);
        }
    }
    check_launch(axis == 0 ? "batch_x" : "batch_y");
}

template <class T, int F, int P, int E>
int batch_variant_resident(int axis, int tile)
{
    if (axis != 0) return resident_blocks((const void*) k_batch_y<T, F, P, E, false>, kYBlock);
    switch (tile) {
    case 128: return resident_blocks((const void*) k_batch_x<T, F, P, E, 128>, 128);
    case 512: return resident_blocks((const void*) k_batch_x<T, F, P, E, 512>, 512);
    default: return resident_blocks((const void*) k_batch_x<T, F, P, E, 256>, 256);
    }
}

// Calls fn.template operator()<F, P, E>() for the variant's compile-time parameters.
template <class Fn>
auto with_variant(const SweepVariant& v, Fn&& fn)
{
    auto by_eos = [&](auto f, auto p) {
        return v.eos == 0 ? fn.template operator()<decltype(f)::value, decltype(p)::value, 0>()
                          : fn.template operator()<decltype(f)::value, decltype(p)::value, 1>();
    };
    auto by_proj = [&](auto f) {
        return v.proj == 2 ? by_eos(f, std::integral_constant<int, 2>{}) : by_eos(f, std::integral_constant<int, 1>{});
    };
    switch (v.flux) {
    case 0: return by_proj(std::integral_constant<int, 0>{});
    case 1: return by_proj(std::integral_constant<int, 1>{});
    case 2: return by_proj(std::integral_constant<int, 2>{});
    default: return by_proj(std::integral_constant<int, 3>{});
    }
}

}  // namespace sweep

template <class T>
void launch_sweep(int axis, const SweepVariant& v, const SweepArgs<T>& a, hipStream_t s)
{
    switch (v.flux) {
    case 0: sweep::launch_proj_dispatch<T, 0>(axis, v, a, s); break;
    case 1: sweep::launch_proj_dispatch<T, 1>(axis, v, a, s); break;
    case 2: sweep::launch_proj_dispatch<T, 2>(axis, v, a, s); break;
    default: sweep::launch_proj_dispatch<T, 3>(axis, v, a, s); break;
    }
}

template <class T>
void launch_batch_sweep(int axis, const SweepVariant& v, const BatchArgs<T>& b, const void* items, int nitems, int grid,
                        int tile, int cpt, hipStream_t s)
{
    sweep::with_variant(v, [&]<int F, int P, int E>() {
        sweep::launch_batch_variant<T, F, P, E>(axis, b, items, nitems, grid, tile, cpt, s);
        return 0;
    });
}

template <class T>
int batch_resident_blocks(int axis, const SweepVariant& v, int tile)
{
    return sweep::with_variant(v,
                               [&]<int F, int P, int E>() { return sweep::batch_variant_resident<T, F, P, E>(axis, tile); });
}

}  // namespace armon
