// Fused Lagrange + remap sweeps: device state and launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace armon {

// Time-stepping state kept on the device between cycles (armon_2D.cpp:229-252, 347-361).
struct DeviceState {
    double current_dt;
    double next_dt;
    double time;
    double error_dt;
    unsigned long long dt_min_bits;  // CFL minimum of the current state (bit pattern of a non-negative double)
    int cycle;
    int active;                      // 1 while the current cycle runs
    int error;                       // 1 after an invalid time step
    int error_cycle;
};

struct TimeStepConfig {
    double cfl, Dt, max_time;
    int max_cycles, cst_dt, dt_on_even_cycles;
};

// Everything a sweep kernel needs. Arrays use the reference layout: row-major with
// nb_ghosts ghost layers on every side, index(i, j) = index_start + j*row_length + i.
template <class T>
struct SweepArgs {
    const T* in_rho; const T* in_u; const T* in_v; const T* in_E;
    T* out_rho; T* out_u; T* out_v; T* out_E;
    const T* stale_p; const T* stale_c;   // pmat/cmat, read only for ghost layers beyond the stencil width
    DeviceState* st;
    int nx, ny, row_length, index_start, stencil_width;
    int seg_len;                          // Y sweep: rows per marching run (set by the launcher)
    int x_slots;                          // X sweep: slots per row, nx + 2 x the ghost layers read
    float x_inv_slots;                    // 1 / x_slots
    int dx_mode, reduce;                  // dx_mode: DX_POW2 (inv_dx = 1 / dx) or DX_RECIP (inv_dx = RN(1 / dx))
    double gamma, dx, inv_dx, dt_factor;
    double lo_u, lo_v, hi_u, hi_v;        // BC factors: Left/Bottom (lo), Right/Top (hi)
    double cfl_dx, cfl_dy;
};

// Compile-time sweep variant: flux = 0 Godunov, 1..3 GAD with limiter None/Minmod/Superbee;
// proj = 1 (Euler) or 2 (Euler 2nd order); eos = 0 perfect gas, 1 Bizarrium.
struct SweepVariant {
    int flux, proj, eos;
};

// Ghost layers read by a sweep on each side: faces from -e to n+e (e = 1 + [Euler 2nd]),
// a GAD face reading one more cell each side (ranges.cpp:14-30, acoustic.cpp:51-63).
inline int read_extent(int flux, int proj) { return 2 + (flux > 0 ? 1 : 0) + (proj == 2 ? 1 : 0); }

template <class T>
void launch_sweep(int axis, const SweepVariant& v, const SweepArgs<T>& a, hipStream_t s);

// Output cells per X-sweep tile of `tile` threads (the single-domain sweep's: 256).
int x_tile_outputs(const SweepVariant& v, int tile = 256);

void launch_time_step(DeviceState* st, const TimeStepConfig& c, hipStream_t s);

// ---- batched ensemble --------------------------------------------------------------------

// One domain of a batch: its slice of the packed arrays and its per-axis constants.
struct BatchDomain {
    long long offset;              // first element of the domain in the packed arrays
    int nx, ny, row_length, index_start;
    double dx[2], inv_dx[2];       // per axis; inv_dx = RN(1 / dx), for the batch's DX_RECIP division
    double cfl_dx, cfl_dy;
    double bc[4][2];               // per side (Left, Right, Top, Bottom): u, v factors
    int x_slots;                   // X sweep: slots per row (nx + 2 x the ghost layers read)
    float x_inv_slots;             // 1 / x_slots
};

// Work items: an X tile (domain; row and slot within that row of the tile's first thread,
// see sweep_x_tile) or a Y run of one wave (domain, first column, output rows [y0, y1)).
struct BatchItemX { int domain, row, slot, pad; };
struct BatchItemY { int domain, x0, y0, y1; };

template <class T>
struct BatchArgs {
    const T* in_rho; const T* in_u; const T* in_v; const T* in_E;
    T* out_rho; T* out_u; T* out_v; T* out_E;
    const T* stale_p; const T* stale_c;
    const BatchDomain* domains;
    DeviceState* states;
    int stencil_width, reduce;
    double gamma, dt_factor;
};

// One launch sweeps every domain of the batch along `axis`: `grid` persistent blocks walk
// the work items (BatchItemX for X, BatchItemY for Y); X tiles are `tile` cells wide (128,
// 256 or 512, the width the X items were cut for), `cpt` cells per thread (1, or 2 for tiles of
// 256 or 512 cells, one block per tile whatever `grid`).
template <class T>
void launch_batch_sweep(int axis, const SweepVariant& v, const BatchArgs<T>& b, const void* items, int nitems,
                        int grid, int tile, int cpt, hipStream_t s);

// Blocks of the batched sweep kernel of this variant (X tiles of `tile` threads) that fit on the
// device at once.
template <class T>
int batch_resident_blocks(int axis, const SweepVariant& v, int tile);

// The time step of every domain (one thread each), per-domain configuration.
void launch_batch_time_step(DeviceState* states, const TimeStepConfig* cfgs, int n, hipStream_t s);

// CFL minimum of the real cells into st->dt_min_bits (after resetting it).
template <class T>
void launch_cfl_real(DeviceState* st, const T* u, const T* v, const T* c, int nx, int ny, int row_length,
                     int index_start, double cfl_dx, double cfl_dy, hipStream_t s);

// Element-wise conversions between the reference FP64 arrays and FP32 working buffers.
void launch_convert_d2f(const double* in, float* out, int64_t n, hipStream_t s);
void launch_convert_f2d(const float* in, double* out, int64_t n, hipStream_t s);

// EOS of a (possibly FP32) state over the real cells into FP64 pmat/cmat/gmat.
template <class T>
void launch_eos_real(const T* rho, const T* u, const T* v, const T* E, double* p, double* c, double* g, int eos,
                     double gamma, int nx, int ny, int row_length, int index_start, hipStream_t s);

}  // namespace armon
