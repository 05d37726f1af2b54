// Fused engine, FP64 instantiation of the sweep kernels (sweep_kernels.hpp), plus the
// device time step, the CFL reduction of a whole state, the EOS of a whole state and the
// FP32 <-> FP64 conversions used by the FP32-mixed form.
#include "sweep_kernels.hpp"

#include <cfloat>
#include <stdexcept>
#include <string>

namespace armon {

using namespace sweep;

namespace {

// ---------------------------------------------------------------------------------------
// Time step (armon_2D.cpp:212-252 and the loop condition of :347)
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ void time_step(DeviceState* st, const TimeStepConfig& c)
{
    if (st->error || !(st->time < c.max_time && st->cycle < c.max_cycles)) {
        st->active = 0;
        return;
    }
    double current = st->next_dt;
    double next;
    if (c.cst_dt) {
        next = c.Dt;
    } else if (!c.dt_on_even_cycles || (st->cycle % 2 == 0) || current == 0) {
        const double dt = bits_to_double(st->dt_min_bits);
        double local;
        if (!finite_d(dt) || dt <= 0) local = dt;
        else if (current == 0) local = c.cfl * dt;
        else local = kmin(c.cfl * dt, 1.05 * current);
        if (!finite_d(local) || local <= 0) {
            st->error = 1;
            st->error_cycle = st->cycle;
            st->error_dt = local;
            st->active = 0;
            return;
        }
        next = local;
    } else {
        next = current;
    }
    if (st->cycle == 0) current = next;
    st->current_dt = current;
    st->next_dt = next;
    st->dt_min_bits = double_to_bits(DBL_MAX);
    st->time += current;
    st->cycle += 1;
    st->active = 1;
}

__global__ void k_time_step(DeviceState* st, TimeStepConfig c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) time_step(st, c);
}

__global__ void k_batch_time_step(DeviceState* states, const TimeStepConfig* cfgs, int n)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n) time_step(states + d, cfgs[d]);
}

__global__ void k_reset_min(DeviceState* st) { st->dt_min_bits = double_to_bits(DBL_MAX); }

template <class T>
__global__ void __launch_bounds__(256) k_cfl_real(DeviceState* st, const T* u, const T* v, const T* c, int nx, int ny,
                                                  int row_length, int index_start, double cfl_dx, double cfl_dy)
{
    double local = DBL_MAX;
    const int64_t total = (int64_t) nx * ny;
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t) gridDim.x * blockDim.x) {
        const int j = (int) (k / nx), i = (int) (k - (int64_t) j * nx);
        const int idx = index_start + j * row_length + i;
        local = kmin(local, cfl_cell<double>(cfl_dx, cfl_dy, (double) u[idx], (double) v[idx], (double) c[idx], 1.0));
    }
    reduce_cfl(st, local);
}

__global__ void k_d2f(const double* in, float* out, int64_t n)
{
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t) gridDim.x * blockDim.x)
        out[k] = (float) in[k];
}

__global__ void k_f2d(const float* in, double* out, int64_t n)
{
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t) gridDim.x * blockDim.x)
        out[k] = (double) in[k];
}

template <class T>
__global__ void __launch_bounds__(256) k_eos_real(const T* rho, const T* u, const T* v, const T* E, double* p,
                                                  double* c, double* g, int eos, double gamma, int nx, int ny,
                                                  int row_length, int index_start)
{
    const int64_t total = (int64_t) nx * ny;
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t) gridDim.x * blockDim.x) {
        const int j = (int) (k / nx), i = (int) (k - (int64_t) j * nx);
        const int idx = index_start + j * row_length + i;
        double pd, cd, gd;
        if (eos == 0) perfect_gas<double>(gamma, (double) rho[idx], (double) E[idx], (double) u[idx], (double) v[idx], pd, cd, gd);
        else bizarrium((double) rho[idx], (double) E[idx], (double) u[idx], (double) v[idx], pd, cd, gd);
        p[idx] = pd;
        c[idx] = cd;
        g[idx] = gd;
    }
}

}  // namespace

template void launch_sweep<double>(int, const SweepVariant&, const SweepArgs<double>&, hipStream_t);
template void launch_batch_sweep<double>(int, const SweepVariant&, const BatchArgs<double>&, const void*, int, int, int,
                                       int, hipStream_t);
template int batch_resident_blocks<double>(int, const SweepVariant&, int);

int x_tile_outputs(const SweepVariant& v, int tile) { return tile - 2 * read_extent(v.flux, v.proj); }

void launch_time_step(DeviceState* st, const TimeStepConfig& c, hipStream_t s)
{
    k_time_step<<<1, 64, 0, s>>>(st, c);
    check_launch("time_step");
}

void launch_batch_time_step(DeviceState* states, const TimeStepConfig* cfgs, int n, hipStream_t s)
{
    k_batch_time_step<<<(unsigned) ((n + 63) / 64), 64, 0, s>>>(states, cfgs, n);
    check_launch("batch_time_step");
}

template <class T>
void launch_cfl_real(DeviceState* st, const T* u, const T* v, const T* c, int nx, int ny, int row_length,
                     int index_start, double cfl_dx, double cfl_dy, hipStream_t s)
{
    k_reset_min<<<1, 1, 0, s>>>(st);
    check_launch("cfl_reset");
    k_cfl_real<T><<<grid_for((int64_t) nx * ny, 256, 2048), 256, 0, s>>>(st, u, v, c, nx, ny, row_length, index_start,
                                                                         cfl_dx, cfl_dy);
    check_launch("cfl_real");
}

template void launch_cfl_real<double>(DeviceState*, const double*, const double*, const double*, int, int, int, int,
                                      double, double, hipStream_t);
template void launch_cfl_real<float>(DeviceState*, const float*, const float*, const float*, int, int, int, int,
                                     double, double, hipStream_t);

void launch_convert_d2f(const double* in, float* out, int64_t n, hipStream_t s)
{
    k_d2f<<<grid_for(n, 256, 8192), 256, 0, s>>>(in, out, n);
    check_launch("convert_d2f");
}

void launch_convert_f2d(const float* in, double* out, int64_t n, hipStream_t s)
{
    k_f2d<<<grid_for(n, 256, 8192), 256, 0, s>>>(in, out, n);
    check_launch("convert_f2d");
}

template <class T>
void launch_eos_real(const T* rho, const T* u, const T* v, const T* E, double* p, double* c, double* g, int eos,
                     double gamma, int nx, int ny, int row_length, int index_start, hipStream_t s)
{
    k_eos_real<T><<<grid_for((int64_t) nx * ny, 256, 8192), 256, 0, s>>>(rho, u, v, E, p, c, g, eos, gamma, nx, ny,
                                                                          row_length, index_start);
    check_launch("eos_real");
}

template void launch_eos_real<double>(const double*, const double*, const double*, const double*, double*, double*,
                                      double*, int, double, int, int, int, int, hipStream_t);
template void launch_eos_real<float>(const float*, const float*, const float*, const float*, double*, double*,
                                     double*, int, double, int, int, int, int, hipStream_t);

}  // namespace armon
