// Host driver of the Armon solver: parameters (reference src/parameters.cpp), iteration
// domains (src/ranges.cpp, src/kernels/indexing.cpp) and the time loop (src/armon_2D.cpp),
// with two engines:
//   staged: the reference's schedule, one launch per stage through the drop-in C-ABI and a
//           blocking dt_CFL per cycle;
//   fused:  one fused sweep kernel per axis, dt/time/cycle on the device, CFL reduction in
//           the last sweep's epilogue, cycles replayed from a captured hipGraph.
#include "internal.hpp"
#include "fused.hpp"
#include "device_math.hpp"
#include "trace.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <limits>
#include <string>
#include <vector>

using namespace armon;

namespace {

// ---------------------------------------------------------------------------------------
// Test cases (test_cases.h:25-121)
// ---------------------------------------------------------------------------------------

double test_default_cfl(int test)
{
    switch (test) {
    case 3: return 0.6;
    case 4: return 0.7;
    default: return 0.95;
    }
}

double test_default_max_time(int test)
{
    switch (test) {
    case 3: return 80e-6;
    case 4: return 1.0;
    default: return 0.20;
    }
}

bool is_nan_bits(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    return (b & 0x7FF0000000000000ull) == 0x7FF0000000000000ull && (b & 0x000FFFFFFFFFFFFFull) != 0;
}

bool is_finite_bits(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    return (b & 0x7FF0000000000000ull) != 0x7FF0000000000000ull;
}

void init_indexing(ArmonParams* p)   // parameters.cpp:52-67
{
    p->row_length = p->nb_ghosts * 2 + p->nx;
    p->col_length = p->nb_ghosts * 2 + p->ny;
    p->nb_cells = p->row_length * p->col_length;
    p->ideb = p->row_length * p->nb_ghosts + p->nb_ghosts;
    p->ifin = p->row_length * (p->ny - 1 + p->nb_ghosts) + p->nb_ghosts + p->nx - 1;
    p->index_start = p->ideb;
    p->idx_row = p->row_length;
    p->idx_col = 1;
}

// indexing.cpp:33-53 (Julia-style ranges, step >= 0)
uint64_t range_length(int64_t start, int64_t step, int64_t stop)
{
    if ((start != stop) && ((step > 0) != (stop > start))) return 0;
    return (uint64_t) ((stop - start) / step + 1);
}

int64_t range_last(int64_t start, int64_t step, int64_t stop)
{
    if (stop == start) return stop;
    return stop - (stop - start) % step;
}

void inflate_dir(DomainRange& d, int axis, int64_t n)   // indexing.cpp:5-17
{
    if (axis == ARMON_AXIS_X) { d.row_start -= d.row_step * n; d.row_end += d.row_step * n; }
    else { d.col_start -= d.col_step * n; d.col_end += d.col_step * n; }
}

void expand_dir(DomainRange& d, int axis, int64_t n)    // indexing.cpp:20-30
{
    if (axis == ARMON_AXIS_X) d.row_end += d.row_step * n;
    else d.col_end += d.col_step * n;
}

std::pair<Range, InnerRange2D> iter2D(const DomainRange& d)
{
    Range r;
    InnerRange2D ir;
    armon_domain_iter2D(&d, &r, &ir);
    return {r, ir};
}

std::pair<Range, InnerRange1D> iter1D(const DomainRange& d)
{
    Range r;
    InnerRange1D ir;
    armon_domain_iter1D(&d, &r, &ir);
    return {r, ir};
}

const char* kFieldNames[ARMON_FIELD_COUNT] = {
    "x", "y", "rho", "umat", "vmat", "Emat", "pmat", "cmat", "gmat", "ustar", "pstar",
    "work_array_1", "work_array_2", "work_array_3", "work_array_4", "domain_mask"};

}  // namespace

// ---------------------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------------------

struct ArmonSolver {
    ArmonParams p{};
    int64_t n = 0;
    double* arr[ARMON_FIELD_COUNT] = {};
    view v[ARMON_FIELD_COUNT] = {};
    std::string labels[ARMON_FIELD_COUNT];
    hipStream_t stream = nullptr;

    // fused engine
    DeviceState* st = nullptr;
    float* f32[10] = {};             // FP32-mixed buffers: A rho,u,v,E | B rho,u,v,E | stale p, c
    int sweeps_per_cycle = 2;
    int64_t cycles_enqueued = 0;
    hipGraphExec_t graph[2] = {};    // graph_chunk cycles from an even (0) or odd (1) cycle
    hipGraph_t graph_def[2] = {};
    bool begun = false;
    double initial_mass = 0, initial_energy = 0;
    std::chrono::steady_clock::time_point t_begin;
    int error_count_at_begin = 0;
    bool dt_error = false;           // an invalid time step stopped the loop (armon_2D.cpp:238-242)
    bool compare_stop = false;       // a per-step comparison found a difference
    StageClock clock;                // per-stage device time (on when verbose < 3) and roctx ranges

    ~ArmonSolver();
};

static int g_error_count = 0;

namespace {

void record_error(const char* where, const std::exception& e)
{
    g_error_count++;
    if (auto* ae = dynamic_cast<const Error*>(&e)) report(where, *ae);
    else report_std(where, e);
}

// Swaps the library stream for the solver's one while the staged engine calls the C-ABI.
struct StreamScope {
    void* saved;
    explicit StreamScope(hipStream_t s) : saved(armon_get_stream()) { armon_set_stream(s); }
    ~StreamScope() { armon_set_stream(saved); }
};

view& V(ArmonSolver* s, int f) { return s->v[f]; }

bool step_checkpoint(ArmonSolver* s, const char* label);

// ---- staged engine (armon_2D.cpp:70-209) --------------------------------------------

void staged_numerical_fluxes(ArmonSolver* s)
{
    const ArmonParams& p = s->p;
    const view& u = p.current_axis == ARMON_AXIS_X ? V(s, ARMON_FIELD_UMAT) : V(s, ARMON_FIELD_VMAT);
    DomainRange d = armon_domain_fluxes(&p);
    auto [range, inner] = iter2D(d);
    if (p.scheme == ARMON_SCHEME_GODUNOV)
        acoustic(range, inner, p.s, V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_PSTAR), V(s, ARMON_FIELD_RHO), u,
                 V(s, ARMON_FIELD_PMAT), V(s, ARMON_FIELD_CMAT));
    else
        acoustic_GAD(range, inner, p.s, p.cycle_dt, p.dx, V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_PSTAR),
                     V(s, ARMON_FIELD_RHO), u, V(s, ARMON_FIELD_PMAT), V(s, ARMON_FIELD_CMAT),
                     static_cast<Limiter>(p.limiter));
}

void staged_update_EOS(ArmonSolver* s)
{
    const ArmonParams& p = s->p;
    DomainRange d = armon_real_domain(&p);
    auto [range, inner] = iter2D(d);
    if (p.test == 3)
        bizarrium_EOS(range, inner, V(s, ARMON_FIELD_RHO), V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT),
                      V(s, ARMON_FIELD_EMAT), V(s, ARMON_FIELD_PMAT), V(s, ARMON_FIELD_CMAT), V(s, ARMON_FIELD_GMAT));
    else
        perfect_gas_EOS(range, inner, 1.4, V(s, ARMON_FIELD_RHO), V(s, ARMON_FIELD_EMAT), V(s, ARMON_FIELD_UMAT),
                        V(s, ARMON_FIELD_VMAT), V(s, ARMON_FIELD_PMAT), V(s, ARMON_FIELD_CMAT),
                        V(s, ARMON_FIELD_GMAT));
}

void staged_cell_update(ArmonSolver* s)
{
    const ArmonParams& p = s->p;
    DomainRange d = armon_domain_cell_update(&p);
    auto [range, inner] = iter2D(d);
    view& u = p.current_axis == ARMON_AXIS_X ? V(s, ARMON_FIELD_UMAT) : V(s, ARMON_FIELD_VMAT);
    cell_update(range, inner, p.s, p.dx, p.cycle_dt, V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_PSTAR),
                V(s, ARMON_FIELD_RHO), u, V(s, ARMON_FIELD_EMAT));
}

void staged_boundary_conditions(ArmonSolver* s)
{
    const ArmonParams& p = s->p;
    const int sides_x[2] = {ARMON_SIDE_LEFT, ARMON_SIDE_RIGHT};
    const int sides_y[2] = {ARMON_SIDE_TOP, ARMON_SIDE_BOTTOM};
    const int* sides = p.current_axis == ARMON_AXIS_X ? sides_x : sides_y;
    for (int k = 0; k < 2; k++) {
        int disp = 0;
        DomainRange d = armon_boundary_conditions_domain(&p, sides[k], &disp);
        double uf, vf;
        boundary_factors(p, sides[k], &uf, &vf);
        Range range;
        InnerRange1D inner;
        armon_domain_direct_iter1D(&d, &range, &inner);
        boundary_conditions(range, inner, p.stencil_width, disp, uf, vf, V(s, ARMON_FIELD_RHO),
                            V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT), V(s, ARMON_FIELD_PMAT),
                            V(s, ARMON_FIELD_CMAT), V(s, ARMON_FIELD_GMAT), V(s, ARMON_FIELD_EMAT));
    }
}

void staged_projection_remap(ArmonSolver* s)
{
    const ArmonParams& p = s->p;
    if (p.projection == ARMON_PROJ_NONE) return;
    view& a_rho = V(s, ARMON_FIELD_WORK1);
    view& a_urho = V(s, ARMON_FIELD_WORK2);
    view& a_vrho = V(s, ARMON_FIELD_WORK3);
    view& a_Erho = V(s, ARMON_FIELD_WORK4);
    {
        DomainRange d = armon_domain_advection(&p);
        auto [range, inner] = iter2D(d);
        if (p.projection == ARMON_PROJ_EULER)
            advection_first_order(range, inner, p.s, p.cycle_dt, V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_RHO),
                                  V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT), V(s, ARMON_FIELD_EMAT), a_rho,
                                  a_urho, a_vrho, a_Erho);
        else
            advection_second_order(range, inner, p.s, p.dx, p.cycle_dt, V(s, ARMON_FIELD_USTAR),
                                   V(s, ARMON_FIELD_RHO), V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT),
                                   V(s, ARMON_FIELD_EMAT), a_rho, a_urho, a_vrho, a_Erho);
    }
    DomainRange d = armon_real_domain(&p);
    auto [range, inner] = iter2D(d);
    euler_projection(range, inner, p.s, p.dx, p.cycle_dt, V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_RHO),
                     V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT), V(s, ARMON_FIELD_EMAT), a_rho, a_urho, a_vrho,
                     a_Erho);
}

// armon_2D.cpp:212-226
double staged_local_time_step(ArmonSolver* s, double prev_dt)
{
    const ArmonParams& p = s->p;
    const double dx = p.domain_size[0] / (double) p.nx;
    const double dy = p.domain_size[1] / (double) p.ny;
    DomainRange d = armon_real_domain(&p);
    auto [range, inner] = iter1D(d);
    const double dt = dt_CFL(range, inner, dx, dy, V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT),
                             V(s, ARMON_FIELD_CMAT), V(s, ARMON_FIELD_DOMAIN_MASK));
    if (!is_finite_bits(dt) || dt <= 0) return dt;
    if (prev_dt == 0) return p.cfl * dt;
    return kmin(p.cfl * dt, 1.05 * prev_dt);
}

}  // namespace

namespace {

// armon_2D.cpp:229-252. Returns false on an invalid time step.
bool staged_time_step(ArmonSolver* s)
{
    ArmonParams& p = s->p;
    p.current_cycle_dt = p.next_cycle_dt;
    if (p.cst_dt) {
        p.next_cycle_dt = p.Dt;
    } else if (!p.dt_on_even_cycles || (p.cycle % 2 == 0) || p.current_cycle_dt == 0) {
        const double local_dt = staged_local_time_step(s, p.current_cycle_dt);
        if (!is_finite_bits(local_dt) || local_dt <= 0) {
            char buf[128];
            snprintf(buf, sizeof buf, "Invalid dt at cycle %d: %f", p.cycle, local_dt);
            g_error_count++;
            s->dt_error = true;
            report("time_step", Error("std::runtime_error", buf));
            return false;
        }
        p.next_cycle_dt = local_dt;
    } else {
        p.next_cycle_dt = p.current_cycle_dt;
    }
    if (p.cycle == 0) p.current_cycle_dt = p.next_cycle_dt;
    return true;
}

// armon_2D.cpp:304-324
bool staged_solver_cycle(ArmonSolver* s)
{
    ArmonParams& p = s->p;
    Stage cycle(s->clock, "cycle", nullptr);
    {
        Stage st(s->clock, "time_step", "dt_CFL");
        if (!staged_time_step(s)) return false;
    }
    if (step_checkpoint(s, "time_step")) return false;
    int axes[3];
    double factors[3];
    const int n = armon_split_axes(&p, axes, factors);
    for (int k = 0; k < n; k++) {
        Stage axis(s->clock, "axis", nullptr);
        armon_update_axis(&p, axes[k]);
        p.cycle_dt = p.current_cycle_dt * factors[k];
        { Stage st(s->clock, "EOS", "EOS"); staged_update_EOS(s); }
        if (step_checkpoint(s, "EOS")) return false;
        { Stage st(s->clock, "BC", "boundary_conditions"); staged_boundary_conditions(s); }
        if (step_checkpoint(s, "boundary_conditions")) return false;
        { Stage st(s->clock, "fluxes", "numerical_fluxes"); staged_numerical_fluxes(s); }
        if (step_checkpoint(s, "numerical_fluxes")) return false;
        { Stage st(s->clock, "update", "cell_update"); staged_cell_update(s); }
        if (step_checkpoint(s, "cell_update")) return false;
        { Stage st(s->clock, "remap", "euler_proj"); staged_projection_remap(s); }
        if (step_checkpoint(s, "projection_remap")) return false;
    }
    return true;
}

void conservation(ArmonSolver* s, double* mass, double* energy)
{
    StreamScope scope(s->stream);
    const ArmonParams& p = s->p;
    DomainRange d = armon_real_domain(&p);
    auto [range, inner] = iter1D(d);
    double m = 0, e = 0;
    conservation_vars(range, inner, p.dx, V(s, ARMON_FIELD_RHO), V(s, ARMON_FIELD_EMAT),
                      V(s, ARMON_FIELD_DOMAIN_MASK), m, e);
    *mass = m;
    *energy = e;
}

// ---- fused engine -----------------------------------------------------------------------

bool fused_supported(const ArmonParams& p, std::string& why)
{
    const int flux = p.scheme == ARMON_SCHEME_GODUNOV ? 0 : 1;
    const int R = read_extent(flux, p.projection);
    const int e = 1 + (p.projection == ARMON_PROJ_EULER_2ND ? 1 : 0);
    if (p.projection == ARMON_PROJ_NONE) { why = "projection none"; return false; }
    if (p.nb_ghosts < R) { why = "nb_ghosts is below the ghost layers read by the stencil"; return false; }
    if (p.stencil_width < e) { why = "stencil width below the cell-update ghost extent"; return false; }
    if (p.nx < R || p.ny < R || p.nx < p.stencil_width || p.ny < p.stencil_width) {
        why = "domain smaller than the stencil";
        return false;
    }
    if ((int64_t) p.nb_cells >= (int64_t) 1 << 31) { why = "more than 2^31 cells"; return false; }
    if (p.nx >= 1 << 23) { why = "rows of 2^23 cells or more"; return false; }   // X tiles' slot arithmetic
    // the sweeps divide by the cell sizes through their reciprocals (DX_RECIP)
    for (double d : {p.domain_size[0] / (double) p.nx, p.domain_size[1] / (double) p.ny})
        if (!(d >= 0x1p-100 && d <= 0x1p100)) { why = "a cell size outside [2^-100, 2^100]"; return false; }
    return true;
}

SweepVariant variant_of(const ArmonParams& p)
{
    SweepVariant v;
    v.flux = p.scheme == ARMON_SCHEME_GODUNOV ? 0 : 1 + p.limiter;
    v.proj = p.projection == ARMON_PROJ_EULER_2ND ? 2 : 1;
    v.eos = p.test == 3 ? 1 : 0;
    return v;
}

bool is_pow2(double x)
{
    if (!(x > 0) || !is_finite_bits(x)) return false;
    int e;
    return std::frexp(x, &e) == 0.5;
}


template <class T>
SweepArgs<T> sweep_args(ArmonSolver* s, int axis, double factor, bool reduce, T* const in[4], T* const out[4],
                        const T* stale_p, const T* stale_c)
{
    const ArmonParams& p = s->p;
    SweepArgs<T> a{};
    a.in_rho = in[0]; a.in_u = in[1]; a.in_v = in[2]; a.in_E = in[3];
    a.out_rho = out[0]; a.out_u = out[1]; a.out_v = out[2]; a.out_E = out[3];
    a.stale_p = stale_p;
    a.stale_c = stale_c;
    a.st = s->st;
    a.nx = p.nx;
    a.ny = p.ny;
    a.row_length = p.row_length;
    a.index_start = p.index_start;
    a.stencil_width = p.stencil_width;
    a.seg_len = p.y_run_rows;   // > 0: rows per run; 0: the launcher chooses
    a.gamma = 1.4;
    a.dx = axis == ARMON_AXIS_X ? p.domain_size[0] / (double) p.nx : p.domain_size[1] / (double) p.ny;
    a.dx_mode = is_pow2(a.dx) ? DX_POW2 : DX_RECIP;
    a.inv_dx = 1.0 / a.dx;
    a.dt_factor = factor;
    double lu, lv, hu, hv;
    boundary_factors(p, axis == ARMON_AXIS_X ? ARMON_SIDE_LEFT : ARMON_SIDE_BOTTOM, &lu, &lv);
    boundary_factors(p, axis == ARMON_AXIS_X ? ARMON_SIDE_RIGHT : ARMON_SIDE_TOP, &hu, &hv);
    a.lo_u = lu; a.lo_v = lv; a.hi_u = hu; a.hi_v = hv;
    a.reduce = reduce ? 1 : 0;
    a.cfl_dx = p.domain_size[0] / (double) p.nx;
    a.cfl_dy = p.domain_size[1] / (double) p.ny;
    return a;
}

// Kernel classes for timing: 0 X sweep, 1 Y sweep, 2 time step.
struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<int> cls;
    hipStream_t s;
    bool on = false;
    void begin()
    {
        if (!on) return;
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        hip_check(hipEventRecord(a, s), "hipEventRecord");
        ev.push_back({a, b});
    }
    void end(int c)
    {
        if (!on) return;
        hip_check(hipEventRecord(ev.back().second, s), "hipEventRecord");
        cls.push_back(c);
    }
};

// Enqueues cycle number `cycle` (time step + sweeps) on the solver stream.
void fused_enqueue_cycle(ArmonSolver* s, int64_t cycle, Timer* timer)
{
    const ArmonParams& p = s->p;
    TimeStepConfig tc{p.cfl, p.Dt, p.max_time, p.max_cycles, p.cst_dt, p.dt_on_even_cycles};
    if (timer) timer->begin();
    launch_time_step(s->st, tc, s->stream);
    if (timer) timer->end(2);

    ArmonParams pc = p;
    pc.cycle = (int) cycle;
    int axes[3];
    double factors[3];
    const int n = armon_split_axes(&pc, axes, factors);
    const SweepVariant var = variant_of(p);
    // buffer parity at the start of this cycle
    int64_t parity = (cycle * s->sweeps_per_cycle) & 1;
    for (int k = 0; k < n; k++) {
        const bool reduce = (k == n - 1);
        if (timer) timer->begin();
        if (p.precision == ARMON_PREC_FP64) {
            double* A[4] = {s->arr[ARMON_FIELD_RHO], s->arr[ARMON_FIELD_UMAT], s->arr[ARMON_FIELD_VMAT],
                            s->arr[ARMON_FIELD_EMAT]};
            double* B[4] = {s->arr[ARMON_FIELD_WORK1], s->arr[ARMON_FIELD_WORK2], s->arr[ARMON_FIELD_WORK3],
                            s->arr[ARMON_FIELD_WORK4]};
            auto a = sweep_args<double>(s, axes[k], factors[k], reduce, parity ? B : A, parity ? A : B,
                                        s->arr[ARMON_FIELD_PMAT], s->arr[ARMON_FIELD_CMAT]);
            launch_sweep<double>(axes[k], var, a, s->stream);
        } else {
            float* A[4] = {s->f32[0], s->f32[1], s->f32[2], s->f32[3]};
            float* B[4] = {s->f32[4], s->f32[5], s->f32[6], s->f32[7]};
            auto a = sweep_args<float>(s, axes[k], factors[k], reduce, parity ? B : A, parity ? A : B, s->f32[8],
                                       s->f32[9]);
            launch_sweep<float>(axes[k], var, a, s->stream);
        }
        if (timer) timer->end(axes[k] == ARMON_AXIS_X ? 0 : 1);
        parity ^= 1;
    }
}

// Captures graph_chunk cycles starting at a cycle of parity `parity`. What a cycle enqueues
// depends on its index only through its parity (the splitting order of parameters.cpp:212-255
// and the ping-pong buffer parity), and everything else (dt, time, cycle count) lives on the
// device, so the graph replays any run of graph_chunk cycles that starts at that parity.
void fused_build_graph(ArmonSolver* s, int parity)
{
    const int chunk = s->p.graph_chunk;
    hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (int c = 0; c < chunk; c++) fused_enqueue_cycle(s, parity + c, nullptr);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void) hipStreamEndCapture(s->stream, &g);
        if (g) (void) hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(s->stream, &s->graph_def[parity]), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&s->graph[parity], s->graph_def[parity], nullptr, nullptr, 0),
              "hipGraphInstantiate");
}

// n more cycles: graph launches of graph_chunk cycles while at least that many remain, then
// the rest one cycle at a time.
void fused_advance(ArmonSolver* s, int64_t n)
{
    const int chunk = s->p.graph_chunk;
    while (n > 0) {
        const int64_t c = s->cycles_enqueued;
        if (chunk > 0 && n >= chunk) {
            const int parity = (int) (c & 1);
            if (!s->graph[parity]) fused_build_graph(s, parity);
            Stage st(s->clock, "cycles_graph", "fused cycles (graph)");
            hip_check(hipGraphLaunch(s->graph[parity], s->stream), "hipGraphLaunch");
            s->cycles_enqueued += chunk;
            n -= chunk;
        } else {
            Stage st(s->clock, "cycle", "fused cycles (launch by launch)");
            fused_enqueue_cycle(s, c, nullptr);
            s->cycles_enqueued++;
            n--;
        }
    }
}

DeviceState read_state(ArmonSolver* s)
{
    DeviceState h{};
    hip_check(hipMemcpyAsync(&h, s->st, sizeof h, hipMemcpyDeviceToHost, s->stream), "read device state");
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    return h;
}

void fused_begin(ArmonSolver* s)
{
    ArmonParams& p = s->p;
    const int64_t n = s->n;
    const size_t bytes = sizeof(double) * (size_t) n;
    if (p.precision == ARMON_PREC_FP64) {
        for (int f = 0; f < 4; f++)
            hip_check(hipMemcpyAsync(s->arr[ARMON_FIELD_WORK1 + f], s->arr[ARMON_FIELD_RHO + (f == 3 ? 3 : f)], bytes,
                                     hipMemcpyDeviceToDevice, s->stream),
                      "copy state");
    } else {
        const int src[4] = {ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT, ARMON_FIELD_EMAT};
        for (int f = 0; f < 4; f++) {
            launch_convert_d2f(s->arr[src[f]], s->f32[f], n, s->stream);
            launch_convert_d2f(s->arr[src[f]], s->f32[4 + f], n, s->stream);
        }
        launch_convert_d2f(s->arr[ARMON_FIELD_PMAT], s->f32[8], n, s->stream);
        launch_convert_d2f(s->arr[ARMON_FIELD_CMAT], s->f32[9], n, s->stream);
    }
    DeviceState h{};
    hip_check(hipMemcpyAsync(s->st, &h, sizeof h, hipMemcpyHostToDevice, s->stream), "init device state");
    launch_cfl_real<double>(s->st, s->arr[ARMON_FIELD_UMAT], s->arr[ARMON_FIELD_VMAT], s->arr[ARMON_FIELD_CMAT], p.nx,
                            p.ny, p.row_length, p.index_start, p.domain_size[0] / (double) p.nx,
                            p.domain_size[1] / (double) p.ny, s->stream);
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    s->cycles_enqueued = 0;
    // both graphs now, so that no later advance pays for a capture
    if (p.graph_chunk > 0)
        for (int parity = 0; parity < 2; parity++)
            if (!s->graph[parity]) fused_build_graph(s, parity);
}

// After the loop: put the final state in rho/umat/vmat/Emat and the EOS of the last sweep's
// input state in pmat/cmat/gmat, as the reference leaves them.
void fused_finish(ArmonSolver* s, const DeviceState& h)
{
    const ArmonParams& p = s->p;
    const int64_t sweeps = (int64_t) h.cycle * s->sweeps_per_cycle;
    if (sweeps == 0) return;
    const int cur = (int) (sweeps & 1);    // buffer holding the final state (0: A, 1: B)
    const int eos = p.test == 3 ? 1 : 0;
    double* P = s->arr[ARMON_FIELD_PMAT];
    double* C = s->arr[ARMON_FIELD_CMAT];
    double* G = s->arr[ARMON_FIELD_GMAT];
    if (p.precision == ARMON_PREC_FP64) {
        double* A[4] = {s->arr[ARMON_FIELD_RHO], s->arr[ARMON_FIELD_UMAT], s->arr[ARMON_FIELD_VMAT],
                        s->arr[ARMON_FIELD_EMAT]};
        double* B[4] = {s->arr[ARMON_FIELD_WORK1], s->arr[ARMON_FIELD_WORK2], s->arr[ARMON_FIELD_WORK3],
                        s->arr[ARMON_FIELD_WORK4]};
        double** in = cur ? A : B;   // input of the last sweep
        launch_eos_real<double>(in[0], in[1], in[2], in[3], P, C, G, eos, 1.4, p.nx, p.ny, p.row_length,
                                p.index_start, s->stream);
        if (cur)
            for (int f = 0; f < 4; f++)
                hip_check(hipMemcpyAsync(A[f], B[f], sizeof(double) * (size_t) s->n, hipMemcpyDeviceToDevice,
                                         s->stream),
                          "copy final state");
    } else {
        float* in = cur ? s->f32[0] : s->f32[4];
        launch_eos_real<float>(in, in == s->f32[0] ? s->f32[1] : s->f32[5], in == s->f32[0] ? s->f32[2] : s->f32[6],
                               in == s->f32[0] ? s->f32[3] : s->f32[7], P, C, G, eos, 1.4, p.nx, p.ny, p.row_length,
                               p.index_start, s->stream);
        const int dst[4] = {ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT, ARMON_FIELD_EMAT};
        for (int f = 0; f < 4; f++) launch_convert_f2d(s->f32[(cur ? 4 : 0) + f], s->arr[dst[f]], s->n, s->stream);
    }
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

// Mass and energy of the fused engine's current state (the buffer written by the last sweep).
void fused_conservation(ArmonSolver* s, const DeviceState& h, double* mass, double* energy)
{
    const ArmonParams& p = s->p;
    const int cur = (int) (((int64_t) h.cycle * s->sweeps_per_cycle) & 1);
    view rho, E;
    std::vector<void*> tmp;
    if (p.precision == ARMON_PREC_FP64) {
        rho = view{cur ? s->arr[ARMON_FIELD_WORK1] : s->arr[ARMON_FIELD_RHO], s->n, nullptr};
        E = view{cur ? s->arr[ARMON_FIELD_WORK4] : s->arr[ARMON_FIELD_EMAT], s->n, nullptr};
    } else {
        // widen the FP32 buffers into the gmat / pstar arrays, which the fused engine does not use
        launch_convert_f2d(s->f32[cur ? 4 : 0], s->arr[ARMON_FIELD_GMAT], s->n, s->stream);
        launch_convert_f2d(s->f32[cur ? 7 : 3], s->arr[ARMON_FIELD_PSTAR], s->n, s->stream);
        rho = view{s->arr[ARMON_FIELD_GMAT], s->n, nullptr};
        E = view{s->arr[ARMON_FIELD_PSTAR], s->n, nullptr};
    }
    StreamScope scope(s->stream);
    DomainRange d = armon_real_domain(&p);
    auto [range, inner] = iter1D(d);
    double m = 0, e = 0;
    conservation_vars(range, inner, p.dx, rho, E, V(s, ARMON_FIELD_DOMAIN_MASK), m, e);
    *mass = m;
    *energy = e;
}

void fill_result(ArmonSolver* s, ArmonResult* out, double loop_time, int status)
{
    const ArmonParams& p = s->p;
    if (!out) return;
    out->status = status;
    out->cycles = p.cycle;
    out->time = p.time;
    out->current_dt = p.current_cycle_dt;
    out->final_dt = p.next_cycle_dt;
    out->loop_time = loop_time;
    out->grind_time = p.cycle > 0 ? loop_time / ((double) p.cycle * p.nx * p.ny) * 1e6 : 0.0;
    out->initial_mass = s->initial_mass;
    out->initial_energy = s->initial_energy;
}

}  // namespace

ArmonSolver::~ArmonSolver()
{
    for (int k = 0; k < 2; k++) {
        if (graph[k]) (void) hipGraphExecDestroy(graph[k]);
        if (graph_def[k]) (void) hipGraphDestroy(graph_def[k]);
    }
    for (auto* a : arr)
        if (a) (void) hipFree(a);
    for (auto* a : f32)
        if (a) (void) hipFree(a);
    if (st) (void) hipFree(st);
    if (stream) (void) hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------------------

extern "C" {

// parameters.h:34-97
ARMON_EXPORT void armon_params_default(ArmonParams* p)
{
    memset(p, 0, sizeof *p);
    p->test = 0;
    p->scheme = ARMON_SCHEME_GAD;
    p->riemann = ARMON_RIEMANN_ACOUSTIC;
    p->projection = ARMON_PROJ_EULER;
    p->limiter = 1;
    p->current_axis = ARMON_AXIS_X;
    p->axis_splitting = ARMON_SPLIT_SEQUENTIAL;
    p->nb_ghosts = 3;
    p->nx = 10;
    p->ny = 10;
    p->stencil_width = -1;
    p->domain_origin[0] = NAN;
    p->domain_origin[1] = NAN;
    p->max_cycles = 100;
    p->verbose = 2;
    p->output_precision = 9;
    p->output_file = "output_cpp";
    p->engine = ARMON_ENGINE_FUSED;
    p->precision = ARMON_PREC_FP64;
    p->graph_chunk = 16;
    p->compare = 0;
    p->comparison_tolerance = 1e-13;
    p->y_run_rows = 0;
}

// parameters.cpp:11-18
ARMON_EXPORT int armon_min_ghosts(const ArmonParams* p)
{
    int m = 1;
    m += p->scheme == ARMON_SCHEME_GAD;
    m += p->projection == ARMON_PROJ_EULER_2ND;
    m += (p->projection == ARMON_PROJ_EULER_2ND) && (p->scheme == ARMON_SCHEME_GAD);
    return m;
}

// parameters.cpp:21-49
ARMON_EXPORT int armon_params_init(ArmonParams* p)
{
    if (p->test < 0 || p->test > 4) return 1;
    if (p->cfl == 0) p->cfl = test_default_cfl(p->test);
    if (p->max_time == 0) p->max_time = test_default_max_time(p->test);
    if (p->domain_size[0] <= 0 || p->domain_size[1] <= 0) {
        p->domain_size[0] = p->test == 4 ? 2 : 1;
        p->domain_size[1] = p->test == 4 ? 2 : 1;
    }
    if (is_nan_bits(p->domain_origin[0]) || is_nan_bits(p->domain_origin[1])) {
        p->domain_origin[0] = p->test == 4 ? -1 : 0;
        p->domain_origin[1] = p->test == 4 ? -1 : 0;
    }
    if (p->stencil_width == -1) p->stencil_width = armon_min_ghosts(p);
    // The indexing fields are ints (as in the reference's Params): refuse a domain whose array
    // length, ghosts included, does not fit one.
    if ((int64_t) (2 * (int64_t) p->nb_ghosts + p->nx) * (2 * (int64_t) p->nb_ghosts + p->ny) > INT32_MAX) {
        fprintf(stderr, "ERROR: the domain (%d x %d cells, %d ghost layers) has more than 2^31 - 1 cells\n", p->nx,
                p->ny, p->nb_ghosts);
        return 1;
    }
    init_indexing(p);
    armon_update_axis(p, p->current_axis);
    return 0;
}

// parameters.cpp:70-109
ARMON_EXPORT int armon_params_check(const ArmonParams* p)
{
    if (p->projection == ARMON_PROJ_NONE) {
        fprintf(stderr, "ERROR: Lagrangian mode is not supported\n");
        return 0;
    }
    if (p->nx <= 0 || p->ny <= 0) {
        fprintf(stderr, "ERROR: One of the dimensions of the domain is 0 or negative (nx=%d, ny=%d)\n", p->nx, p->ny);
        return 0;
    }
    if (p->cst_dt && p->Dt == 0.) {
        fprintf(stderr, "ERROR: Constant time step is set ('--cst-dt 1') but the initial time step is 0\n");
        return 0;
    }
    if (p->write_output && p->output_file == nullptr) {
        fprintf(stderr, "ERROR: Write output is on but no output file was given\n");
        return 0;
    }
    const int mg = armon_min_ghosts(p);
    if (p->nb_ghosts < mg)
        fprintf(stderr, "WARNING: There is not enough ghost cells for the given parameters. Expected at least %d, got %d\n",
                mg, p->nb_ghosts);
    if (p->stencil_width < mg)
        fprintf(stderr, "WARNING: The stencil width given does not cover the estimated width. Expected at least %d, got %d\n",
                mg, p->stencil_width);
    if (p->stencil_width > p->nb_ghosts) {
        fprintf(stderr, "ERROR: The stencil width cannot be bigger than the number of ghost cells: %d > %d\n",
                p->stencil_width, p->nb_ghosts);
        return 0;
    }
    if (p->y_run_rows < 0) {
        fprintf(stderr, "ERROR: y_run_rows must be positive or 0 (automatic), got %d\n", p->y_run_rows);
        return 0;
    }
    if (p->engine != ARMON_ENGINE_STAGED && p->engine != ARMON_ENGINE_FUSED) {
        fprintf(stderr, "ERROR: unknown engine %d\n", p->engine);
        return 0;
    }
    if (p->compare && p->engine != ARMON_ENGINE_STAGED) {
        fprintf(stderr, "ERROR: the per-step comparison needs the staged engine\n");
        return 0;
    }
    if (p->compare && p->output_file == nullptr) {
        fprintf(stderr, "ERROR: the per-step comparison needs an output file name (the step files' prefix)\n");
        return 0;
    }
    if (p->write_output && p->write_ghosts && p->engine == ARMON_ENGINE_FUSED) {
        // the fused sweeps synthesise ghost cells in registers and LDS and never store them, so
        // the arrays' ghost layers keep their initial values (the reference writes them at every
        // boundary_conditions call, halo_exchange.cpp:23-30)
        fprintf(stderr, "ERROR: writing the ghost cells needs the staged engine (the fused engine keeps none)\n");
        return 0;
    }
    if (p->engine == ARMON_ENGINE_STAGED && p->precision != ARMON_PREC_FP64) {
        fprintf(stderr, "ERROR: the staged engine runs in FP64 only\n");
        return 0;
    }
    if (p->engine == ARMON_ENGINE_FUSED) {
        std::string why;
        if (!fused_supported(*p, why)) {
            fprintf(stderr, "ERROR: the fused engine does not support this configuration (%s); use the staged engine\n",
                    why.c_str());
            return 0;
        }
    }
    return 1;
}

// parameters.cpp:195-210
ARMON_EXPORT void armon_update_axis(ArmonParams* p, int axis)
{
    p->current_axis = axis;
    if (axis == ARMON_AXIS_X) {
        p->s = 1;
        p->dx = p->domain_size[0] / (double) p->nx;
    } else {
        p->s = p->row_length;
        p->dx = p->domain_size[1] / (double) p->ny;
    }
}

// parameters.cpp:212-255
ARMON_EXPORT int armon_split_axes(const ArmonParams* p, int axes[3], double factors[3])
{
    int a1, a2;
    if (p->cycle % 2 == 0) { a1 = ARMON_AXIS_Y; a2 = ARMON_AXIS_X; }
    else { a1 = ARMON_AXIS_X; a2 = ARMON_AXIS_Y; }
    switch (p->axis_splitting) {
    case ARMON_SPLIT_SEQUENTIAL:
        axes[0] = ARMON_AXIS_X; factors[0] = 1.0; axes[1] = ARMON_AXIS_Y; factors[1] = 1.0; return 2;
    case ARMON_SPLIT_SEQUENTIAL_SYM:
        axes[0] = a1; factors[0] = 1.0; axes[1] = a2; factors[1] = 1.0; return 2;
    case ARMON_SPLIT_STRANG:
        axes[0] = a1; factors[0] = 0.5; axes[1] = a2; factors[1] = 1.0; axes[2] = a1; factors[2] = 0.5; return 3;
    case ARMON_SPLIT_X_ONLY:
        axes[0] = ARMON_AXIS_X; factors[0] = 1.0; return 1;
    case ARMON_SPLIT_Y_ONLY:
        axes[0] = ARMON_AXIS_Y; factors[0] = 1.0; return 1;
    default:
        return 0;
    }
}

// test_cases.h:36-38 and the per-test boundaryCondition overrides
ARMON_EXPORT void armon_test_boundary_condition(int test, int side, double* uf, double* vf)
{
    const bool lr = side == ARMON_SIDE_LEFT || side == ARMON_SIDE_RIGHT;
    double u = 1, v = 1;
    switch (test) {
    case 0: if (lr) u = -1; break;                 // Sod
    case 1: if (!lr) v = -1; break;                // Sod_y
    case 2: if (lr) u = -1; else v = -1; break;    // Sod_circ
    case 3:                                        // Bizarrium
        if (side == ARMON_SIDE_LEFT) u = -1;
        else if (side != ARMON_SIDE_RIGHT) v = -1;
        break;
    default: break;                                // Sedov: free flow
    }
    *uf = u;
    *vf = v;
}

// test_cases.h:118-120 via armon_2D.cpp:116-118
ARMON_EXPORT double armon_test_option(const ArmonParams* p)
{
    if (p->test != 4) return 0.;
    const double dx = p->domain_size[0] / (double) p->nx;
    const double dy = p->domain_size[1] / (double) p->ny;
    return std::hypot(dx, dy) / M_SQRT2;
}

ARMON_EXPORT uint64_t armon_domain_col_length(const DomainRange* d) { return range_length(d->col_start, d->col_step, d->col_end); }
ARMON_EXPORT uint64_t armon_domain_row_length(const DomainRange* d) { return range_length(d->row_start, d->row_step, d->row_end); }
ARMON_EXPORT uint64_t armon_domain_length(const DomainRange* d) { return armon_domain_col_length(d) * armon_domain_row_length(d); }
ARMON_EXPORT int64_t armon_domain_begin(const DomainRange* d) { return d->col_start + d->row_start; }
ARMON_EXPORT int64_t armon_domain_end(const DomainRange* d)
{
    return range_last(d->col_start, d->col_step, d->col_end) + range_last(d->row_start, d->row_step, d->row_end);
}

// indexing.cpp:83-89
ARMON_EXPORT void armon_domain_iter1D(const DomainRange* d, Range* r, InnerRange1D* ir)
{
    r->start = 0;
    r->end = armon_domain_end(d) - armon_domain_begin(d) + 1;
    ir->start = armon_domain_begin(d);
    ir->step = d->row_step;
}

// indexing.cpp:74-80
ARMON_EXPORT void armon_domain_direct_iter1D(const DomainRange* d, Range* r, InnerRange1D* ir)
{
    r->start = 0;
    r->end = (Idx) armon_domain_length(d);
    ir->start = armon_domain_begin(d);
    ir->step = d->row_step;
}

// indexing.cpp:92-98
ARMON_EXPORT void armon_domain_iter2D(const DomainRange* d, Range* r, InnerRange2D* ir)
{
    r->start = 0;
    r->end = (Idx) armon_domain_length(d);
    ir->main_range_start = d->col_start;
    ir->main_range_step = d->col_step;
    ir->row_range_start = d->row_start;
    ir->row_range_length = armon_domain_row_length(d);
}

// ranges.h:44-47
ARMON_EXPORT int64_t armon_index_1D(const ArmonParams* p, int i, int j)
{
    return (int64_t) p->index_start + (int64_t) j * p->idx_row + (int64_t) i * p->idx_col;
}

// ranges.cpp:5-11
ARMON_EXPORT DomainRange armon_real_domain(const ArmonParams* p)
{
    return DomainRange{armon_index_1D(p, 0, 0), p->row_length, armon_index_1D(p, 0, p->ny - 1), 0, 1, p->nx - 1};
}

// ranges.cpp:14-30
ARMON_EXPORT DomainRange armon_domain_fluxes(const ArmonParams* p)
{
    DomainRange d = armon_real_domain(p);
    inflate_dir(d, p->current_axis, 1 + (p->projection == ARMON_PROJ_EULER_2ND ? 1 : 0));
    expand_dir(d, p->current_axis, 1);
    return d;
}

// ranges.cpp:33-46
ARMON_EXPORT DomainRange armon_domain_cell_update(const ArmonParams* p)
{
    DomainRange d = armon_real_domain(p);
    inflate_dir(d, p->current_axis, 1 + (p->projection == ARMON_PROJ_EULER_2ND ? 1 : 0));
    return d;
}

// ranges.cpp:49-54
ARMON_EXPORT DomainRange armon_domain_advection(const ArmonParams* p)
{
    DomainRange d = armon_real_domain(p);
    expand_dir(d, p->current_axis, 1);
    return d;
}

// ranges.cpp:57-63
ARMON_EXPORT DomainRange armon_complete_domain(const ArmonParams* p)
{
    DomainRange d = armon_real_domain(p);
    inflate_dir(d, ARMON_AXIS_X, p->nb_ghosts);
    inflate_dir(d, ARMON_AXIS_Y, p->nb_ghosts);
    return d;
}

// ranges.cpp:66-106
ARMON_EXPORT DomainRange armon_boundary_conditions_domain(const ArmonParams* p, int side, int* disp)
{
    int64_t i_start, stride, loop_range;
    switch (side) {
    case ARMON_SIDE_LEFT:
        stride = p->row_length; i_start = armon_index_1D(p, -1, 0); loop_range = p->ny; *disp = 1; break;
    case ARMON_SIDE_RIGHT:
        stride = p->row_length; i_start = armon_index_1D(p, p->nx, 0); loop_range = p->ny; *disp = -1; break;
    case ARMON_SIDE_TOP:
        stride = 1; i_start = armon_index_1D(p, 0, p->ny); loop_range = p->nx; *disp = -p->row_length; break;
    default:
        stride = 1; i_start = armon_index_1D(p, 0, -1); loop_range = p->nx; *disp = p->row_length; break;
    }
    return DomainRange{0, 1, 0, i_start, stride, i_start + loop_range * stride - stride};
}

// ---- solver -------------------------------------------------------------------------

ARMON_EXPORT ArmonSolver* armon_solver_create(const ArmonParams* params)
{
    auto* s = new ArmonSolver();
    try {
        s->p = *params;
        if (!armon_params_check(&s->p)) throw invalid_argument("invalid parameters (armon_params_check failed)");
        s->n = s->p.nb_cells;
        const size_t bytes = sizeof(double) * (size_t) s->n;
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        s->clock.set_stream(s->stream);
        s->clock.enable(s->p.verbose < 3);
        for (int f = 0; f < ARMON_FIELD_COUNT; f++) {
            void* ptr = nullptr;
            hip_check(hipMalloc(&ptr, bytes), "hipMalloc(field)");
            s->arr[f] = static_cast<double*>(ptr);
            s->labels[f] = std::string("Armon_") + kFieldNames[f];
            s->v[f] = view{s->arr[f], s->n, nullptr};
        }
        for (int f = 0; f < ARMON_FIELD_COUNT; f++) s->v[f].label = s->labels[f].c_str();
        // the reference allocates without initialisation; start from NaN so that a read of a
        // never-written cell cannot go unnoticed
        for (int f = 0; f < ARMON_FIELD_COUNT; f++)
            hip_check(hipMemsetAsync(s->arr[f], 0xFF, bytes, s->stream), "hipMemset");
        void* st = nullptr;
        hip_check(hipMalloc(&st, sizeof(DeviceState)), "hipMalloc(state)");
        s->st = static_cast<DeviceState*>(st);
        if (s->p.engine == ARMON_ENGINE_FUSED && s->p.precision == ARMON_PREC_FP32_MIXED) {
            for (int k = 0; k < 10; k++) {
                void* ptr = nullptr;
                hip_check(hipMalloc(&ptr, sizeof(float) * (size_t) s->n), "hipMalloc(fp32)");
                s->f32[k] = static_cast<float*>(ptr);
            }
        }
        int axes[3];
        double factors[3];
        s->sweeps_per_cycle = armon_split_axes(&s->p, axes, factors);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        return s;
    } catch (const std::exception& e) {
        record_error("armon_solver_create", e);
        delete s;
        return nullptr;
    }
}

ARMON_EXPORT void armon_solver_destroy(ArmonSolver* s) { delete s; }

ARMON_EXPORT int armon_solver_get_params(const ArmonSolver* s, ArmonParams* out)
{
    *out = s->p;
    return 0;
}

ARMON_EXPORT void* armon_solver_stream(ArmonSolver* s) { return s->stream; }

// armon_2D.cpp:112-127
ARMON_EXPORT int armon_solver_init_test(ArmonSolver* s, bool debug_indexes)
{
    try {
        StreamScope scope(s->stream);
        const int e0 = g_error_count;
        const ArmonParams& p = s->p;
        DomainRange d = armon_complete_domain(&p);
        auto [range, inner] = iter2D(d);
        const double test_option = armon_test_option(&p);
        Stage st(s->clock, "init_test", "init_test");
        init_test(range, inner, p.row_length, p.nb_ghosts, p.nx, p.ny, p.domain_size[0], p.domain_size[1],
                  p.domain_origin[0], p.domain_origin[1], 0, 0, p.nx, p.ny, V(s, ARMON_FIELD_X), V(s, ARMON_FIELD_Y),
                  V(s, ARMON_FIELD_RHO), V(s, ARMON_FIELD_EMAT), V(s, ARMON_FIELD_UMAT), V(s, ARMON_FIELD_VMAT),
                  V(s, ARMON_FIELD_DOMAIN_MASK), V(s, ARMON_FIELD_PMAT), V(s, ARMON_FIELD_CMAT),
                  V(s, ARMON_FIELD_GMAT), V(s, ARMON_FIELD_USTAR), V(s, ARMON_FIELD_PSTAR),
                  static_cast<Test>(p.test), debug_indexes, test_option);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        return g_error_count != e0 ? 1 : 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_init_test", e);
        return 1;
    }
}

// armon_2D.cpp:327-345 (before the loop)
ARMON_EXPORT int armon_solver_begin(ArmonSolver* s)
{
    try {
        StreamScope scope(s->stream);
        ArmonParams& p = s->p;
        s->error_count_at_begin = g_error_count;
        s->dt_error = false;
        p.cycle = 0;
        p.time = 0.;
        p.current_cycle_dt = 0.;
        p.next_cycle_dt = 0.;
        s->t_begin = std::chrono::steady_clock::now();
        s->compare_stop = false;
        armon_update_axis(&p, ARMON_AXIS_X);
        { Stage st(s->clock, "EOS_init", nullptr); staged_update_EOS(s); }
        step_checkpoint(s, "EOS_init");
        if (p.verbose <= 1) conservation(s, &s->initial_mass, &s->initial_energy);
        if (p.engine == ARMON_ENGINE_FUSED) fused_begin(s);
        s->begun = true;
        return g_error_count != s->error_count_at_begin ? 1 : 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_begin", e);
        return 1;
    }
}

ARMON_EXPORT int armon_solver_advance(ArmonSolver* s, int n)
{
    try {
        if (!s->begun) throw invalid_argument("armon_solver_begin must be called first");
        StreamScope scope(s->stream);
        ArmonParams& p = s->p;
        if (p.engine == ARMON_ENGINE_FUSED) {
            fused_advance(s, n);
            return 0;
        }
        if (s->compare_stop) return 1;
        for (int k = 0; k < n && p.time < p.max_time && p.cycle < p.max_cycles; k++) {
            const int e0 = g_error_count;
            if (!staged_solver_cycle(s) || g_error_count != e0) return 1;
            s->clock.collect(false);   // the stages that have completed (dt_CFL waited for the cycle's start)
            if (p.verbose <= 1) {
                double m, e;
                conservation(s, &m, &e);
                printf("Cycle = %4d, dt = %.18f, t = %.18f, |dM| = %#8.6g%%, |dE| = %#8.6g%%\n", p.cycle + 1,
                       p.current_cycle_dt, p.time, std::abs(s->initial_mass - m) / s->initial_mass * 100,
                       std::abs(s->initial_energy - e) / s->initial_energy * 100);
            }
            p.time += p.current_cycle_dt;
            p.cycle++;
        }
        return 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_advance", e);
        return 1;
    }
}

ARMON_EXPORT int armon_solver_finish(ArmonSolver* s, ArmonResult* out)
{
    try {
        ArmonParams& p = s->p;
        int status = 0;
        if (p.engine == ARMON_ENGINE_FUSED) {
            const DeviceState h = read_state(s);
            p.cycle = h.cycle;
            p.time = h.time;
            p.current_cycle_dt = h.current_dt;
            p.next_cycle_dt = h.next_dt;
            if (h.cycle > 0) {
                // leave current_axis/dx as the last sweep of the last cycle set them
                ArmonParams last = p;
                last.cycle = h.cycle - 1;
                int axes[3];
                double factors[3];
                const int n = armon_split_axes(&last, axes, factors);
                armon_update_axis(&p, axes[n - 1]);
                p.cycle_dt = p.current_cycle_dt * factors[n - 1];
            }
            fused_finish(s, h);
            if (h.error) {
                char buf[128];
                snprintf(buf, sizeof buf, "Invalid dt at cycle %d: %f", h.error_cycle, h.error_dt);
                g_error_count++;
                report("time_step", Error("std::runtime_error", buf));
                status = 1;
            }
        } else {
            hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
            if (s->dt_error) status = 1;
            else if (s->compare_stop) status = 3;
        }
        const double loop_time =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t_begin).count();
        s->clock.collect(true);
        if (g_error_count != s->error_count_at_begin && status == 0) status = 2;
        fill_result(s, out, loop_time, status);
        if (out && p.verbose <= 1) conservation(s, &out->final_mass, &out->final_energy);
        return status;
    } catch (const std::exception& e) {
        record_error("armon_solver_finish", e);
        fill_result(s, out, 0.0, 2);
        return 2;
    }
}

// armon_2D.cpp:327-386
ARMON_EXPORT int armon_solver_time_loop(ArmonSolver* s, ArmonResult* out)
{
    if (armon_solver_begin(s)) {
        fill_result(s, out, 0.0, 2);
        return 2;
    }
    ArmonParams& p = s->p;
    if (p.engine == ARMON_ENGINE_FUSED) {
        try {
            const int chunk = p.verbose <= 1 ? 1 : (p.graph_chunk > 0 ? p.graph_chunk : 16);
            double prev_time = 0.0;
            while (true) {
                fused_advance(s, chunk);
                const DeviceState h = read_state(s);
                if (p.verbose <= 1 && h.active) {
                    // armon_2D.cpp:351-357; the state is in the buffer of the last sweep
                    double m, e;
                    fused_conservation(s, h, &m, &e);
                    printf("Cycle = %4d, dt = %.18f, t = %.18f, |dM| = %#8.6g%%, |dE| = %#8.6g%%\n", h.cycle,
                           h.current_dt, prev_time, std::abs(s->initial_mass - m) / s->initial_mass * 100,
                           std::abs(s->initial_energy - e) / s->initial_energy * 100);
                    prev_time = h.time;
                }
                if (!h.active || h.error) break;
            }
        } catch (const std::exception& e) {
            record_error("armon_solver_time_loop", e);
            fill_result(s, out, 0.0, 2);
            return 2;
        }
    } else {
        const int e0 = g_error_count;
        while (p.time < p.max_time && p.cycle < p.max_cycles) {
            if (armon_solver_advance(s, 1) || g_error_count != e0) break;
        }
    }
    return armon_solver_finish(s, out);
}

ARMON_EXPORT int armon_solver_stage_count(ArmonSolver* s)
{
    s->clock.collect(true);
    return (int) s->clock.table().size();
}

ARMON_EXPORT int armon_solver_stage_time(ArmonSolver* s, int k, const char** label, double* ms, int64_t* calls)
{
    if (k < 0 || k >= (int) s->clock.table().size()) return 1;
    auto it = s->clock.table().begin();
    std::advance(it, k);
    if (label) *label = it->first.c_str();
    if (ms) *ms = it->second.ms;
    if (calls) *calls = it->second.calls;
    return 0;
}

ARMON_EXPORT int armon_solver_kernel_times(ArmonSolver* s, int n, ArmonKernelTimes* out)
{
    try {
        if (!s->begun) throw invalid_argument("armon_solver_begin must be called first");
        if (s->p.engine != ARMON_ENGINE_FUSED) throw invalid_argument("kernel timing is for the fused engine");
        memset(out, 0, sizeof *out);
        Timer t;
        t.s = s->stream;
        t.on = true;
        for (int k = 0; k < n; k++) {
            fused_enqueue_cycle(s, s->cycles_enqueued, &t);
            s->cycles_enqueued++;
        }
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        for (size_t k = 0; k < t.ev.size(); k++) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, t.ev[k].first, t.ev[k].second), "hipEventElapsedTime");
            out->launches[t.cls[k]]++;
            out->total_ms[t.cls[k]] += ms;
            (void) hipEventDestroy(t.ev[k].first);
            (void) hipEventDestroy(t.ev[k].second);
        }
        return 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_kernel_times", e);
        return 1;
    }
}

ARMON_EXPORT int armon_solver_read_field(ArmonSolver* s, int field, double* host, int64_t n)
{
    try {
        if (field < 0 || field >= ARMON_FIELD_COUNT) throw out_of_range("unknown field");
        if (n != s->n) throw invalid_argument("host buffer size must equal nb_cells");
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(host, s->arr[field], sizeof(double) * (size_t) n, hipMemcpyDeviceToHost), "read field");
        return 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_read_field", e);
        return 1;
    }
}

ARMON_EXPORT int armon_solver_write_field(ArmonSolver* s, int field, const double* host, int64_t n)
{
    try {
        if (field < 0 || field >= ARMON_FIELD_COUNT) throw out_of_range("unknown field");
        if (n != s->n) throw invalid_argument("host buffer size must equal nb_cells");
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(s->arr[field], host, sizeof(double) * (size_t) n, hipMemcpyHostToDevice), "write field");
        return 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_write_field", e);
        return 1;
    }
}

ARMON_EXPORT int armon_solver_conservation(ArmonSolver* s, double* mass, double* energy)
{
    try {
        conservation(s, mass, energy);
        return 0;
    } catch (const std::exception& e) {
        record_error("armon_solver_conservation", e);
        return 1;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Output and the full run (io.cpp:10-44, armon_2D.cpp:389-431)
// ---------------------------------------------------------------------------------------

namespace {

// The main variables of the output and comparison files (data.h:120-130).
const int kMainVars[6] = {ARMON_FIELD_X, ARMON_FIELD_Y, ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT,
                          ARMON_FIELD_PMAT};
using HostVars = std::vector<std::vector<double>>;

HostVars read_main_vars(ArmonSolver* s)
{
    HostVars h(6, std::vector<double>((size_t) s->n));
    for (int f = 0; f < 6; f++)
        if (armon_solver_read_field(s, kMainVars[f], h[f].data(), s->n)) throw hip_failure("reading a field failed");
    return h;
}

// Cells of the files: the real domain, plus the ghost layers with write_ghosts (io.cpp:12-15).
struct FileCells {
    int i_deb, i_fin, j_deb, j_fin;
    explicit FileCells(const ArmonParams& p)
        : i_deb(p.write_ghosts ? -p.nb_ghosts : 0), i_fin(p.nx + (p.write_ghosts ? p.nb_ghosts : 0)),
          j_deb(p.write_ghosts ? -p.nb_ghosts : 0), j_fin(p.ny + (p.write_ghosts ? p.nb_ghosts : 0))
    {
    }
};

// io.cpp:10-44
void write_output_host(const ArmonParams& p, const HostVars& h, const char* file_name)
{
    std::ofstream file(file_name);
    const FileCells c(p);
    const int j_deb = c.j_deb, j_fin = c.j_fin, i_deb = c.i_deb, i_fin = c.i_fin;
    file << std::scientific;
    file.precision(p.output_precision);
    const int width = p.output_precision + 7;
    for (int j = j_deb; j < j_fin; j++) {
        for (int i = i_deb; i < i_fin; i++) {
            const int64_t idx = armon_index_1D(&p, i, j);
            file << std::setw(width) << h[0][idx];
            for (int f = 1; f < 6; f++) file << ", " << std::setw(width) << h[f][idx];
            file << "\n";
        }
        file << "\n";
    }
    if (p.verbose < 2) printf("Wrote to file: %s\n", file_name);
}

void write_output(ArmonSolver* s, const char* file_name) { write_output_host(s->p, read_main_vars(s), file_name); }

// io.h:14-17
bool is_approx(double a, double b, double tol) { return std::abs(a - b) <= tol * std::max(std::abs(a), std::abs(b)); }

// io.cpp:55-74: values separated by ',' and whitespace, one cell per line, in file order.
void load_data(const ArmonParams& p, HostVars& d, std::istream& file)
{
    const FileCells c(p);
    for (int j = c.j_deb; j < c.j_fin; j++)
        for (int i = c.i_deb; i < c.i_fin; i++) {
            const int64_t idx = armon_index_1D(&p, i, j);
            file >> d[0][idx];
            for (int f = 1; f < 6; f++) {
                file.ignore(1, ',');
                file >> d[f][idx];
            }
        }
}

// io.cpp:77-132: number of cells differing beyond the tolerance, reported row by row.
int compare_with_reference(const ArmonParams& p, const HostVars& ref, const HostVars& d)
{
    static const char* names[6] = {"x", "y", "rho", "umat", "vmat", "pmat"};
    const FileCells c(p);
    int total = 0;
    for (int j = c.j_deb; j < c.j_fin; j++) {
        const int64_t row_deb = armon_index_1D(&p, c.i_deb, j);
        const int64_t row_fin = armon_index_1D(&p, c.i_fin, j);
        for (int f = 0; f < 6; f++) {
            int count = 0;
            int64_t diff_start = row_fin, diff_end = row_deb;
            double max_diff = 0;
            long max_ulp = 0;
            for (int64_t idx = row_deb; idx < row_fin; idx++) {
                const double rv = ref[f][idx], v = d[f][idx];
                if (is_approx(rv, v, p.comparison_tolerance)) continue;
                count++;
                diff_start = std::min(idx, diff_start);
                diff_end = std::max(idx, diff_end);
                const double diff = std::abs(rv - v);
                if (diff > max_diff) {
                    max_diff = diff;
                    max_ulp = (long) (max_diff / std::abs(rv * std::numeric_limits<double>::epsilon()));
                }
            }
            total += count;
            if (count > 0)
                printf("Row %3d has %3d differences in '%4s' from %3lld to %3lld, max diff: %.15g (%ld ulp)\n", j + 1,
                       count, names[f], (long long) (diff_start - row_deb + 1), (long long) (diff_end - row_deb + 1),
                       max_diff, max_ulp);
        }
    }
    return total;
}

// io.cpp:135-146
bool compare_with_file(const ArmonParams& p, const HostVars& d, const std::string& name)
{
    std::ifstream file(name);
    file.exceptions(std::ios::failbit | std::istream::badbit);
    HostVars ref(6, std::vector<double>(d[0].size(), 0.0));
    load_data(p, ref, file);
    return compare_with_reference(p, ref, d) > 0;
}

// io.cpp:149-171
bool compare_time_step(const ArmonParams& p, const std::string& name)
{
    std::ifstream file(name);
    file.exceptions(std::ios::failbit | std::istream::badbit);
    double ref_dt;
    file >> ref_dt;
    const bool different = !is_approx(ref_dt, p.current_cycle_dt, p.comparison_tolerance);
    if (different) {
        const double diff = ref_dt - p.current_cycle_dt;
        printf("Time step difference: ref \xce\x94t = %.15g, \xce\x94t = %.15g, diff = %.15g (%ld ulp)\n", ref_dt,
               p.current_cycle_dt, diff, (long) (diff / std::numeric_limits<double>::epsilon()));
    }
    return different;
}

}  // namespace

namespace {

// armon_2D.cpp:255-284: with `compare`, the state after each step is compared with the file
// <output_file>_<cycle>_<step>_<axis>; on a difference the state is written to <file>_diff
// and the run stops there.
bool step_checkpoint(ArmonSolver* s, const char* label)
{
    const ArmonParams& p = s->p;
    if (!p.compare) return false;
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    char buf[100];
    snprintf(buf, sizeof buf, "_%03d_%s", p.cycle, label);
    const std::string name = std::string(p.output_file) + buf + "_" + (p.current_axis == ARMON_AXIS_X ? "X" : "Y");
    HostVars h = read_main_vars(s);
    bool different;
    try {
        different = strcmp(label, "time_step") == 0 ? compare_time_step(p, name) : compare_with_file(p, h, name);
    } catch (const std::exception& e) {
        fprintf(stderr, "Error while comparing with file '%s': %s\n", name.c_str(), e.what());
        different = true;
    }
    if (different) {
        const std::string diff = name + "_diff";
        write_output_host(p, h, diff.c_str());
        printf("Difference file written to %s\n", diff.c_str());
        s->compare_stop = true;
    }
    return different;
}

}  // namespace

extern "C" ARMON_EXPORT int armon_run(ArmonParams* p, ArmonResult* out)
{
    ArmonResult local{};
    if (!out) out = &local;
    ArmonSolver* s = armon_solver_create(p);
    if (!s) return 2;
    int status = armon_solver_init_test(s, false);
    if (status == 0) {
        try {
            if (step_checkpoint(s, "init_test")) status = 3;   // armon_2D.cpp:401
        } catch (const std::exception& e) {
            record_error("armon_run", e);
            status = 2;
        }
        if (status != 0) {
            out->status = status;
            armon_solver_destroy(s);
            return status;
        }
    }
    if (status == 0) status = armon_solver_time_loop(s, out);
    if (status == 0 && p->verbose < 4) {
        printf("\n");
        printf("Time:       %.4g seconds\n", out->loop_time);
        printf("Grind time: %.4g \xc2\xb5s/cell/cycle\n", out->grind_time);
        printf("Cells/sec:  %.4g Mega cells/sec\n", out->grind_time > 0 ? 1. / out->grind_time : 0.0);
        printf("Cycles:     %d\n", out->cycles);
        printf("Final dt:   %.18f\n\n", out->final_dt);
    }
    if (status == 0 && p->verbose < 3) {
        // armon_2D.cpp:420-428, with device time per stage
        double total = 0;
        for (const auto& [label, e] : s->clock.table()) total += e.ms;
        printf("Total time for each step:\n");
        for (const auto& [label, e] : s->clock.table())
            printf(" - %-32s %10.5f ms (%5.2f%%)\n", label.c_str(), e.ms, total > 0 ? e.ms / total * 100 : 0.0);
    }
    if (status == 0 && p->write_output) write_output(s, p->output_file);
    if (status == 0 && p->write_throughput) {
        if (FILE* f = fopen("cell_throughput", "w")) {
            fprintf(f, "%f", out->grind_time > 0 ? 1. / out->grind_time : 0.0);
            fclose(f);
        }
    }
    armon_solver_get_params(s, p);
    armon_solver_destroy(s);
    return status;
}
