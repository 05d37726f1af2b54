// Batched ensemble: many independent Armon domains of different sizes advanced together.
//
// The state of all domains is packed into one allocation per field (each domain keeps the
// reference layout of its own size, with its ghost layers, at an aligned offset). Every
// sweep of a cycle is one launch of a persistent kernel over a work list covering all
// domains: X items are row tiles, Y items are 64-column runs of rows, one per wave. Each
// domain has its own DeviceState (dt, time, cycle, end flag) and TimeStepConfig, so the time
// step of every domain is computed on the device by one small kernel per cycle, and a
// domain that has reached its end time simply stops being updated while the others go on.
// Cycles are replayed from a captured hipGraph. A domain's arithmetic is that of the
// single-domain fused sweeps, so its result is the single-domain result.
#include "internal.hpp"
#include "fused.hpp"
#include "device_math.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

using namespace armon;

namespace {

constexpr int64_t kAlign = 64;   // domain slices start on 64-element boundaries
constexpr int kWave = 64;
constexpr int kYWavesPerBlock = 4;
constexpr int kBatchTileWidths[3] = {128, 256, 512};   // X tile widths the autotuner tries

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (double) (splitmix64(s) >> 11) * 0x1.0p-53; }

struct StreamScope {
    void* saved;
    explicit StreamScope(hipStream_t s) : saved(armon_get_stream()) { armon_set_stream(s); }
    ~StreamScope() { armon_set_stream(saved); }
};

template <class T>
T* device_alloc(size_t count, const char* what)
{
    void* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(T) * (count > 0 ? count : 1)), what);
    return static_cast<T*>(p);
}

}  // namespace

struct ArmonBatch {
    ArmonParams common{};
    std::vector<ArmonParams> p;          // per domain, through armon_params_init
    std::vector<ArmonTestState> states;
    std::vector<int64_t> offset;         // first element of each domain in the packed arrays
    int64_t total = 0;                   // packed elements
    int64_t real_cells = 0;
    double* arr[ARMON_FIELD_COUNT] = {};
    float* f32[10] = {};                 // FP32-mixed: A rho,u,v,E | B rho,u,v,E | stale p, c
    BatchDomain* d_domains = nullptr;
    DeviceState* d_states = nullptr;
    DeviceState* d_tune = nullptr;       // scratch states used while timing candidates
    TimeStepConfig* d_cfgs = nullptr;
    std::vector<BatchItemX> items_x;
    std::vector<BatchItemY> items_y;
    BatchItemX* d_items_x = nullptr;
    BatchItemY* d_items_y = nullptr;
    ArmonBatchTuning tuning{};
    SweepVariant var{};
    int sweeps_per_cycle = 2;
    int64_t cycles_enqueued = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[2] = {};        // graph_chunk cycles from an even (0) or odd (1) cycle
    hipGraph_t graph_def[2] = {};
    bool initialized = false, begun = false, finished = false;
    std::vector<DeviceState> final_states;

    ~ArmonBatch();
    view slice(int field, int d)
    {
        return view{arr[field] + offset[d], (int64_t) p[d].nb_cells, nullptr};
    }
    bool fp64() const { return common.precision == ARMON_PREC_FP64; }
};

ArmonBatch::~ArmonBatch()
{
    for (int k = 0; k < 2; k++) {
        if (graph[k]) (void) hipGraphExecDestroy(graph[k]);
        if (graph_def[k]) (void) hipGraphDestroy(graph_def[k]);
    }
    for (auto* a : arr)
        if (a) (void) hipFree(a);
    for (auto* a : f32)
        if (a) (void) hipFree(a);
    for (void* a : {(void*) d_domains, (void*) d_states, (void*) d_tune, (void*) d_cfgs, (void*) d_items_x,
                    (void*) d_items_y})
        if (a) (void) hipFree(a);
    if (stream) (void) hipStreamDestroy(stream);
}

namespace {

void fail(const char* where, const std::exception& e)
{
    if (auto* ae = dynamic_cast<const Error*>(&e)) report(where, *ae);
    else report_std(where, e);
}

// Work lists. X: every tile of every domain's stream of padded rows (the tile mapping of
// sweep_x_tile: tiles of W outputs overlapping by 2R slots). Y: every (64-column group, run of
// at most `run` rows) of every domain, band by band, so that the waves of a block march
// neighbouring column groups of one band (measured 2% faster in FP64 and 12% in FP32-mixed
// than column group by column group); then longest runs first (a stable sort) so that a
// grid-stride walk over them balances well.
// Destroys the captured graphs: they hold the work lists' addresses and the launch grids.
void drop_graphs(ArmonBatch* b)
{
    for (int k = 0; k < 2; k++) {
        if (b->graph[k]) (void) hipGraphExecDestroy(b->graph[k]);
        if (b->graph_def[k]) (void) hipGraphDestroy(b->graph_def[k]);
        b->graph[k] = nullptr;
        b->graph_def[k] = nullptr;
    }
}

void build_items(ArmonBatch* b, int run)
{
    drop_graphs(b);
    const int W = x_tile_outputs(b->var, b->tuning.tile_x);
    const int R = read_extent(b->var.flux, b->var.proj);
    b->items_x.clear();
    b->items_y.clear();
    const int n = (int) b->p.size();
    for (int d = 0; d < n; d++) {
        const ArmonParams& q = b->p[d];
        const int64_t slots = q.nx + 2 * R;
        const int64_t tiles = (slots * q.ny - 2 * R + W - 1) / W;
        for (int64_t t = 0; t < tiles; t++) {
            const int64_t s0 = t * W;
            b->items_x.push_back(BatchItemX{d, (int) (s0 / slots), (int) (s0 % slots), 0});
        }
        for (int y0 = 0; y0 < q.ny; y0 += run)
            for (int x0 = 0; x0 < q.nx; x0 += kWave) b->items_y.push_back(BatchItemY{d, x0, y0, std::min(y0 + run, q.ny)});
    }
    std::stable_sort(b->items_y.begin(), b->items_y.end(),
                     [](const BatchItemY& a, const BatchItemY& c) { return a.y1 - a.y0 > c.y1 - c.y0; });
    if (b->d_items_x) (void) hipFree(b->d_items_x);
    if (b->d_items_y) (void) hipFree(b->d_items_y);
    b->d_items_x = nullptr;
    b->d_items_y = nullptr;
    b->d_items_x = device_alloc<BatchItemX>(b->items_x.size(), "hipMalloc(items_x)");
    b->d_items_y = device_alloc<BatchItemY>(b->items_y.size(), "hipMalloc(items_y)");
    hip_check(hipMemcpy(b->d_items_x, b->items_x.data(), sizeof(BatchItemX) * b->items_x.size(), hipMemcpyHostToDevice),
              "upload items_x");
    hip_check(hipMemcpy(b->d_items_y, b->items_y.data(), sizeof(BatchItemY) * b->items_y.size(), hipMemcpyHostToDevice),
              "upload items_y");
    b->tuning.y_run_rows = run;
    b->tuning.items_x = (int) b->items_x.size();
    b->tuning.items_y = (int) b->items_y.size();
}

// Blocks of a launch: `grid` when positive and below one block per item, else one per item.
int blocks_x(const ArmonBatch* b, int grid)
{
    const int all = (int) b->items_x.size();
    return grid > 0 && grid < all ? grid : std::max(all, 1);
}

int blocks_y(const ArmonBatch* b, int grid)
{
    const int all = (int) ((b->items_y.size() + kYWavesPerBlock - 1) / kYWavesPerBlock);
    return grid > 0 && grid < all ? grid : std::max(all, 1);
}

template <class T>
BatchArgs<T> batch_args(ArmonBatch* b, int parity, double factor, bool reduce, DeviceState* states)
{
    BatchArgs<T> a{};
    T* A[4];
    T* B[4];
    const T* sp;
    const T* sc;
    if constexpr (std::is_same_v<T, double>) {
        A[0] = b->arr[ARMON_FIELD_RHO]; A[1] = b->arr[ARMON_FIELD_UMAT];
        A[2] = b->arr[ARMON_FIELD_VMAT]; A[3] = b->arr[ARMON_FIELD_EMAT];
        B[0] = b->arr[ARMON_FIELD_WORK1]; B[1] = b->arr[ARMON_FIELD_WORK2];
        B[2] = b->arr[ARMON_FIELD_WORK3]; B[3] = b->arr[ARMON_FIELD_WORK4];
        sp = b->arr[ARMON_FIELD_PMAT];
        sc = b->arr[ARMON_FIELD_CMAT];
    } else {
        for (int f = 0; f < 4; f++) { A[f] = b->f32[f]; B[f] = b->f32[4 + f]; }
        sp = b->f32[8];
        sc = b->f32[9];
    }
    T** in = parity ? B : A;
    T** out = parity ? A : B;
    a.in_rho = in[0]; a.in_u = in[1]; a.in_v = in[2]; a.in_E = in[3];
    a.out_rho = out[0]; a.out_u = out[1]; a.out_v = out[2]; a.out_E = out[3];
    a.stale_p = sp;
    a.stale_c = sc;
    a.domains = b->d_domains;
    a.states = states;
    a.stencil_width = b->common.stencil_width;
    a.reduce = reduce ? 1 : 0;
    a.gamma = 1.4;
    a.dt_factor = factor;
    return a;
}

void launch_one_sweep(ArmonBatch* b, int axis, int parity, double factor, bool reduce, DeviceState* states)
{
    const bool x = axis == ARMON_AXIS_X;
    const void* items = x ? (const void*) b->d_items_x : (const void*) b->d_items_y;
    const int nitems = x ? (int) b->items_x.size() : (int) b->items_y.size();
    const int grid = x ? blocks_x(b, b->tuning.grid_x) : blocks_y(b, b->tuning.grid_y);
    if (b->fp64())
        launch_batch_sweep<double>(axis, b->var, batch_args<double>(b, parity, factor, reduce, states), items, nitems,
                                   grid, b->tuning.tile_x, b->tuning.x_cells_per_thread, b->stream);
    else
        launch_batch_sweep<float>(axis, b->var, batch_args<float>(b, parity, factor, reduce, states), items, nitems,
                                  grid, b->tuning.tile_x, b->tuning.x_cells_per_thread, b->stream);
}

struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<int> cls;
    hipStream_t s = nullptr;
    void begin()
    {
        hipEvent_t a, c;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&c), "hipEventCreate");
        hip_check(hipEventRecord(a, s), "hipEventRecord");
        ev.push_back({a, c});
    }
    void end(int k)
    {
        hip_check(hipEventRecord(ev.back().second, s), "hipEventRecord");
        cls.push_back(k);
    }
    ~Timer()
    {
        for (auto& e : ev) {
            (void) hipEventDestroy(e.first);
            (void) hipEventDestroy(e.second);
        }
    }
};

// Cycle `cycle` of every domain: the per-domain time step, then the sweeps of the splitting.
void enqueue_cycle(ArmonBatch* b, int64_t cycle, Timer* t)
{
    if (t) t->begin();
    launch_batch_time_step(b->d_states, b->d_cfgs, (int) b->p.size(), b->stream);
    if (t) t->end(2);
    ArmonParams pc = b->common;
    pc.cycle = (int) cycle;
    int axes[3];
    double factors[3];
    const int n = armon_split_axes(&pc, axes, factors);
    int parity = (int) ((cycle * b->sweeps_per_cycle) & 1);
    for (int k = 0; k < n; k++) {
        if (t) t->begin();
        launch_one_sweep(b, axes[k], parity, factors[k], k == n - 1, b->d_states);
        if (t) t->end(axes[k] == ARMON_AXIS_X ? 0 : 1);
        parity ^= 1;
    }
}

// graph_chunk cycles from a cycle of parity `parity`: a cycle's launches depend on its index
// only through its parity (splitting order, buffer parity), so the graph replays any run of
// graph_chunk cycles that starts at that parity.
void build_graph(ArmonBatch* b, int parity)
{
    hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (int c = 0; c < b->common.graph_chunk; c++) enqueue_cycle(b, parity + c, nullptr);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void) hipStreamEndCapture(b->stream, &g);
        if (g) (void) hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(b->stream, &b->graph_def[parity]), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&b->graph[parity], b->graph_def[parity], nullptr, nullptr, 0),
              "hipGraphInstantiate");
}

// n more cycles: graph launches while at least graph_chunk cycles remain, then one by one.
void advance(ArmonBatch* b, int64_t n)
{
    const int chunk = b->common.graph_chunk;
    while (n > 0) {
        const int64_t c = b->cycles_enqueued;
        if (chunk > 0 && n >= chunk) {
            const int parity = (int) (c & 1);
            if (!b->graph[parity]) build_graph(b, parity);
            hip_check(hipGraphLaunch(b->graph[parity], b->stream), "hipGraphLaunch");
            b->cycles_enqueued += chunk;
            n -= chunk;
        } else {
            enqueue_cycle(b, c, nullptr);
            b->cycles_enqueued++;
            n--;
        }
    }
}

std::vector<DeviceState> read_states(ArmonBatch* b)
{
    std::vector<DeviceState> h(b->p.size());
    hip_check(hipMemcpyAsync(h.data(), b->d_states, sizeof(DeviceState) * h.size(), hipMemcpyDeviceToHost, b->stream),
              "read device states");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    return h;
}

// Working buffers from the reference arrays: FP64 copies A into B (ghost layers the sweeps
// never write must hold the initial values in both); FP32-mixed rounds A into both FP32
// buffers and pmat/cmat into the stale-ghost buffers.
void prepare_buffers(ArmonBatch* b)
{
    const size_t n = (size_t) b->total;
    const int src[4] = {ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT, ARMON_FIELD_EMAT};
    if (b->fp64()) {
        for (int f = 0; f < 4; f++)
            hip_check(hipMemcpyAsync(b->arr[ARMON_FIELD_WORK1 + f], b->arr[src[f]], sizeof(double) * n,
                                     hipMemcpyDeviceToDevice, b->stream),
                      "copy state");
    } else {
        for (int f = 0; f < 4; f++) {
            launch_convert_d2f(b->arr[src[f]], b->f32[f], (int64_t) n, b->stream);
            launch_convert_d2f(b->arr[src[f]], b->f32[4 + f], (int64_t) n, b->stream);
        }
        launch_convert_d2f(b->arr[ARMON_FIELD_PMAT], b->f32[8], (int64_t) n, b->stream);
        launch_convert_d2f(b->arr[ARMON_FIELD_CMAT], b->f32[9], (int64_t) n, b->stream);
    }
}

// Average time of `reps` launches of one sweep under the current tuning (after one warm-up).
double time_sweep(ArmonBatch* b, int axis, int reps)
{
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    float ms = 0;
    try {
        const bool reduce = axis == ARMON_AXIS_Y;
        launch_one_sweep(b, axis, 0, 1.0, reduce, b->d_tune);
        hip_check(hipEventRecord(e0, b->stream), "hipEventRecord");
        for (int r = 0; r < reps; r++) launch_one_sweep(b, axis, 0, 1.0, reduce, b->d_tune);
        hip_check(hipEventRecord(e1, b->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    } catch (...) {
        (void) hipEventDestroy(e0);
        (void) hipEventDestroy(e1);
        throw;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    return ms / reps;
}

void check_ready(const ArmonBatch* b, bool need_init, bool not_running)
{
    if (need_init && !b->initialized) throw invalid_argument("armon_batch_init must be called first");
    if (not_running && b->begun && !b->finished) throw invalid_argument("the batch is running (finish it first)");
}

}  // namespace

extern "C" {

ARMON_EXPORT int armon_batch_random_specs(uint64_t seed, int n, int64_t min_area, int64_t max_area, double max_aspect,
                                          ArmonDomainSpec* out)
{
    if (n < 0 || !out || min_area < 256 || max_area < min_area || !(max_aspect >= 1.0)) return 1;
    uint64_t s = seed;
    const int tests[3] = {0, 2, 4};   // Sod, Sod_circ, Sedov
    const double la = std::log((double) min_area), lb = std::log((double) max_area);
    const double lr = std::log(max_aspect);
    for (int k = 0; k < n; k++) {
        ArmonDomainSpec& d = out[k];
        memset(&d, 0, sizeof d);
        const double area = std::exp(la + (lb - la) * uniform01(s));
        const double aspect = std::exp(lr * uniform01(s));
        const bool wide = uniform01(s) < 0.5;
        const int longer = std::max(16, (int) std::lround(std::sqrt(area * aspect)));
        const int shorter = std::max(16, (int) std::lround(std::sqrt(area / aspect)));
        d.nx = wide ? longer : shorter;
        d.ny = wide ? shorter : longer;
        d.test = tests[std::min(2, (int) (3.0 * uniform01(s)))];
        d.max_time = 0.0;
        ArmonParams q;
        armon_params_default(&q);
        q.test = d.test;
        q.nx = d.nx;
        q.ny = d.ny;
        armon_params_init(&q);
        armon_reference_test_state(d.test, armon_test_option(&q), &d.state);
        const double f_rho_hi = 0.8 + 0.4 * uniform01(s), f_p_hi = 0.8 + 0.4 * uniform01(s);
        const double f_rho_lo = 0.8 + 0.4 * uniform01(s), f_p_lo = 0.8 + 0.4 * uniform01(s);
        // p = (gamma - 1) rho (E - |u|^2 / 2) with u = 0 initially: scaling E by f_p / f_rho scales p by f_p
        d.state.high_rho *= f_rho_hi;
        d.state.high_E *= f_p_hi / f_rho_hi;
        d.state.low_rho *= f_rho_lo;
        d.state.low_E *= f_p_lo / f_rho_lo;
    }
    return 0;
}

ARMON_EXPORT ArmonBatch* armon_batch_create(const ArmonParams* common, int n, const ArmonDomainSpec* specs)
{
    auto* b = new ArmonBatch();
    try {
        if (n <= 0 || !specs) throw invalid_argument("a batch needs at least one domain");
        b->common = *common;
        if (b->common.engine != ARMON_ENGINE_FUSED) throw invalid_argument("the batch runs the fused engine");
        if (b->common.precision != ARMON_PREC_FP64 && b->common.precision != ARMON_PREC_FP32_MIXED)
            throw invalid_argument("unknown precision");
        const bool bizarrium = specs[0].test == 3;
        b->p.resize(n);
        b->states.resize(n);
        b->offset.resize(n);
        int64_t off = 0;
        for (int d = 0; d < n; d++) {
            const ArmonDomainSpec& sp = specs[d];
            if (sp.test < 0 || sp.test > 4) throw invalid_argument("unknown test in domain " + std::to_string(d));
            if ((sp.test == 3) != bizarrium)
                throw invalid_argument("the domains of a batch share their EOS: Bizarrium for all or for none");
            ArmonParams q = b->common;
            q.test = sp.test;
            q.nx = sp.nx;
            q.ny = sp.ny;
            q.max_time = sp.max_time > 0 ? sp.max_time : 0.0;   // 0: the test's default
            q.cfl = common->cfl > 0 ? common->cfl : 0.0;             // 0: the test's default
            q.domain_size[0] = q.domain_size[1] = 0;     // the test's domain
            q.domain_origin[0] = q.domain_origin[1] = NAN;
            q.cycle = 0;
            q.verbose = 5;
            q.write_output = 0;
            if (armon_params_init(&q)) throw invalid_argument("armon_params_init failed for domain " + std::to_string(d));
            if (!armon_params_check(&q)) throw invalid_argument("invalid parameters for domain " + std::to_string(d));
            b->p[d] = q;
            b->states[d] = sp.state;
            b->offset[d] = off;
            off += ((int64_t) q.nb_cells + kAlign - 1) / kAlign * kAlign;
            b->real_cells += (int64_t) q.nx * q.ny;
        }
        b->total = off;
        // shared parameters as every domain sees them (stencil width resolved by params_init)
        b->common.stencil_width = b->p[0].stencil_width;
        b->common.test = b->p[0].test;
        b->var.flux = b->common.scheme == ARMON_SCHEME_GODUNOV ? 0 : 1 + b->common.limiter;
        b->var.proj = b->common.projection == ARMON_PROJ_EULER_2ND ? 2 : 1;
        b->var.eos = bizarrium ? 1 : 0;
        int axes[3];
        double factors[3];
        b->sweeps_per_cycle = armon_split_axes(&b->common, axes, factors);

        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        const size_t bytes = sizeof(double) * (size_t) b->total;
        for (int f = 0; f < ARMON_FIELD_COUNT; f++) {
            b->arr[f] = device_alloc<double>((size_t) b->total, "hipMalloc(field)");
            // NaN everywhere, padding included: a read of a never-written cell cannot go unnoticed
            hip_check(hipMemsetAsync(b->arr[f], 0xFF, bytes, b->stream), "hipMemset");
        }
        if (!b->fp64())
            for (int k = 0; k < 10; k++) b->f32[k] = device_alloc<float>((size_t) b->total, "hipMalloc(fp32)");

        std::vector<BatchDomain> dom(n);
        std::vector<TimeStepConfig> cfg(n);
        const int sides[4] = {ARMON_SIDE_LEFT, ARMON_SIDE_RIGHT, ARMON_SIDE_TOP, ARMON_SIDE_BOTTOM};
        for (int d = 0; d < n; d++) {
            const ArmonParams& q = b->p[d];
            BatchDomain& D = dom[d];
            memset(&D, 0, sizeof D);
            D.offset = b->offset[d];
            D.nx = q.nx;
            D.ny = q.ny;
            D.row_length = q.row_length;
            D.index_start = q.index_start;
            D.dx[0] = q.domain_size[0] / (double) q.nx;
            D.dx[1] = q.domain_size[1] / (double) q.ny;
            for (int a = 0; a < 2; a++) {
                // the batched sweeps divide by dx through its reciprocal (DX_RECIP)
                if (!(D.dx[a] >= 0x1p-100 && D.dx[a] <= 0x1p100))
                    throw invalid_argument("domain " + std::to_string(d) + ": cell size outside [2^-100, 2^100]");
                D.inv_dx[a] = 1.0 / D.dx[a];
            }
            D.cfl_dx = D.dx[0];
            D.cfl_dy = D.dx[1];
            for (int k = 0; k < 4; k++) boundary_factors(q, sides[k], &D.bc[k][0], &D.bc[k][1]);
            D.x_slots = q.nx + 2 * read_extent(b->var.flux, b->var.proj);
            D.x_inv_slots = 1.0f / (float) D.x_slots;
            cfg[d] = TimeStepConfig{q.cfl, q.Dt, q.max_time, q.max_cycles, q.cst_dt, q.dt_on_even_cycles};
        }
        b->d_domains = device_alloc<BatchDomain>(n, "hipMalloc(domains)");
        b->d_cfgs = device_alloc<TimeStepConfig>(n, "hipMalloc(configs)");
        b->d_states = device_alloc<DeviceState>(n, "hipMalloc(states)");
        b->d_tune = device_alloc<DeviceState>(n, "hipMalloc(tune states)");
        hip_check(hipMemcpy(b->d_domains, dom.data(), sizeof(BatchDomain) * n, hipMemcpyHostToDevice), "upload domains");
        hip_check(hipMemcpy(b->d_cfgs, cfg.data(), sizeof(TimeStepConfig) * n, hipMemcpyHostToDevice), "upload configs");
        // default configuration: 256-cell X tiles of one cell per thread, 256-row Y runs, one
        // block per work item
        b->tuning.tile_x = 256;
        b->tuning.x_cells_per_thread = 1;
        b->tuning.grid_x = 0;
        b->tuning.grid_y = 0;
        build_items(b, 256);
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        return b;
    } catch (const std::exception& e) {
        fail("armon_batch_create", e);
        delete b;
        return nullptr;
    }
}

ARMON_EXPORT void armon_batch_destroy(ArmonBatch* b) { delete b; }
ARMON_EXPORT int armon_batch_count(const ArmonBatch* b) { return (int) b->p.size(); }
ARMON_EXPORT int64_t armon_batch_cells(const ArmonBatch* b) { return b->real_cells; }
ARMON_EXPORT void* armon_batch_stream(ArmonBatch* b) { return b->stream; }

ARMON_EXPORT int armon_batch_domain_params(const ArmonBatch* b, int d, ArmonParams* out)
{
    if (d < 0 || d >= (int) b->p.size()) return 1;
    *out = b->p[d];
    return 0;
}

ARMON_EXPORT int armon_batch_init(ArmonBatch* b)
{
    try {
        check_ready(b, false, true);
        StreamScope scope(b->stream);
        const uint64_t e0 = report_count();
        for (int d = 0; d < (int) b->p.size(); d++) {
            const ArmonParams& q = b->p[d];
            DomainRange dr = armon_complete_domain(&q);
            Range range;
            InnerRange2D inner;
            armon_domain_iter2D(&dr, &range, &inner);
            view x = b->slice(ARMON_FIELD_X, d), y = b->slice(ARMON_FIELD_Y, d), rho = b->slice(ARMON_FIELD_RHO, d),
                 E = b->slice(ARMON_FIELD_EMAT, d), u = b->slice(ARMON_FIELD_UMAT, d), v = b->slice(ARMON_FIELD_VMAT, d),
                 mask = b->slice(ARMON_FIELD_DOMAIN_MASK, d), pm = b->slice(ARMON_FIELD_PMAT, d),
                 cm = b->slice(ARMON_FIELD_CMAT, d), gm = b->slice(ARMON_FIELD_GMAT, d),
                 us = b->slice(ARMON_FIELD_USTAR, d), ps = b->slice(ARMON_FIELD_PSTAR, d);
            init_test_states(range, inner, q.row_length, q.nb_ghosts, q.nx, q.ny, q.domain_size[0], q.domain_size[1],
                             q.domain_origin[0], q.domain_origin[1], 0, 0, q.nx, q.ny, x, y, rho, E, u, v, mask, pm, cm,
                             gm, us, ps, static_cast<Test>(q.test), armon_test_option(&q), b->states[d]);
        }
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        if (report_count() != e0) return 1;
        b->initialized = true;
        b->begun = false;
        b->finished = false;
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_init", e);
        return 1;
    }
}

ARMON_EXPORT int armon_batch_get_tuning(const ArmonBatch* b, ArmonBatchTuning* out)
{
    *out = b->tuning;
    return 0;
}

ARMON_EXPORT int armon_batch_set_tuning(ArmonBatch* b, const ArmonBatchTuning* t)
{
    try {
        check_ready(b, false, true);
        if (t->y_run_rows < 1) throw invalid_argument("y_run_rows must be positive");
        if (t->tile_x != 0 && t->tile_x != 128 && t->tile_x != 256 && t->tile_x != 512)
            throw invalid_argument("tile_x must be 128, 256 or 512 (0: 256)");
        const int tile = t->tile_x ? t->tile_x : 256;
        const int cpt = t->x_cells_per_thread ? t->x_cells_per_thread : 1;
        if (cpt != 1 && cpt != 2) throw invalid_argument("x_cells_per_thread must be 1 or 2 (0: 1)");
        if (cpt == 2 && (tile == 128 || t->grid_x > 0))
            throw invalid_argument("2 cells per thread needs X tiles of 256 or 512 cells and one block per tile "
                                   "(grid_x <= 0)");
        b->tuning.tile_x = tile;
        b->tuning.x_cells_per_thread = cpt;
        b->tuning.grid_x = std::max(0, t->grid_x);
        b->tuning.grid_y = std::max(0, t->grid_y);
        build_items(b, t->y_run_rows);
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_set_tuning", e);
        return 1;
    }
}

ARMON_EXPORT int armon_batch_autotune(ArmonBatch* b, ArmonBatchTuning* out)
{
    try {
        check_ready(b, true, true);
        prepare_buffers(b);
        // every domain active with a small positive dt: the work of a real cycle
        std::vector<DeviceState> tune(b->p.size());
        for (auto& s : tune) {
            memset(&s, 0, sizeof s);
            s.active = 1;
            s.current_dt = 1e-7;
            s.next_dt = 1e-7;
            s.dt_min_bits = 0x7FEFFFFFFFFFFFFFull;   // DBL_MAX
        }
        hip_check(hipMemcpyAsync(b->d_tune, tune.data(), sizeof(DeviceState) * tune.size(), hipMemcpyHostToDevice,
                                 b->stream),
                  "upload tune states");
        const int ry = b->fp64() ? batch_resident_blocks<double>(ARMON_AXIS_Y, b->var, 256)
                                 : batch_resident_blocks<float>(ARMON_AXIS_Y, b->var, 256);
        const int reps = 5;

        // X: tile width (128, 256, 512 cells, one per thread) x (persistent grids of 1, 2, 4
        // resident rounds | one block per tile), then tiles of 256 and 512 cells at two cells
        // per thread (one block per tile)
        double best_x = 1e300;
        int grid_x = 0, tile_x = 256, cpt_x = 1;
        const int run0 = b->tuning.y_run_rows;
        b->tuning.x_cells_per_thread = 1;
        for (int tile : kBatchTileWidths) {
            b->tuning.tile_x = tile;
            build_items(b, run0);
            const int rx = b->fp64() ? batch_resident_blocks<double>(ARMON_AXIS_X, b->var, tile)
                                     : batch_resident_blocks<float>(ARMON_AXIS_X, b->var, tile);
            for (int g : {rx, 2 * rx, 4 * rx, 0}) {
                if (g < 0) continue;
                b->tuning.grid_x = g;
                const double ms = time_sweep(b, ARMON_AXIS_X, reps);
                if (ms < best_x) { best_x = ms; grid_x = g; tile_x = tile; }
            }
        }
        b->tuning.x_cells_per_thread = 2;
        b->tuning.grid_x = 0;
        for (int tile : {256, 512}) {
            b->tuning.tile_x = tile;
            build_items(b, run0);
            const double ms = time_sweep(b, ARMON_AXIS_X, reps);
            if (ms < best_x) { best_x = ms; grid_x = 0; tile_x = tile; cpt_x = 2; }
        }
        b->tuning.tile_x = tile_x;
        b->tuning.x_cells_per_thread = cpt_x;
        b->tuning.grid_x = grid_x;

        // Y: run length x (one resident round of blocks | one block per 4 runs)
        double best_y = 1e300;
        int grid_y = 0, run_y = 256;
        for (int run : {16, 32, 64, 128, 256, 512, 1024}) {
            build_items(b, run);
            for (int g : {ry, 0}) {
                b->tuning.grid_y = g;
                const double ms = time_sweep(b, ARMON_AXIS_Y, reps);
                if (ms < best_y) { best_y = ms; grid_y = g; run_y = run; }
            }
        }
        b->tuning.grid_y = grid_y;
        build_items(b, run_y);
        b->tuning.ms_x = best_x;
        b->tuning.ms_y = best_y;
        if (out) *out = b->tuning;
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_autotune", e);
        return 1;
    }
}

// armon_2D.cpp:327-345 for every domain: EOS of the initial state, then the CFL bound that
// the first time step uses.
ARMON_EXPORT int armon_batch_begin(ArmonBatch* b)
{
    try {
        check_ready(b, true, true);
        StreamScope scope(b->stream);
        const uint64_t e0 = report_count();
        const int n = (int) b->p.size();
        for (int d = 0; d < n; d++) {
            const ArmonParams& q = b->p[d];
            DomainRange dr = armon_real_domain(&q);
            Range range;
            InnerRange2D inner;
            armon_domain_iter2D(&dr, &range, &inner);
            view rho = b->slice(ARMON_FIELD_RHO, d), u = b->slice(ARMON_FIELD_UMAT, d), v = b->slice(ARMON_FIELD_VMAT, d),
                 E = b->slice(ARMON_FIELD_EMAT, d), pm = b->slice(ARMON_FIELD_PMAT, d), cm = b->slice(ARMON_FIELD_CMAT, d),
                 gm = b->slice(ARMON_FIELD_GMAT, d);
            if (q.test == 3) bizarrium_EOS(range, inner, rho, u, v, E, pm, cm, gm);
            else perfect_gas_EOS(range, inner, 1.4, rho, E, u, v, pm, cm, gm);
        }
        if (report_count() != e0) return 1;
        prepare_buffers(b);
        std::vector<DeviceState> zero(n);
        memset(zero.data(), 0, sizeof(DeviceState) * n);
        hip_check(hipMemcpyAsync(b->d_states, zero.data(), sizeof(DeviceState) * n, hipMemcpyHostToDevice, b->stream),
                  "init device states");
        for (int d = 0; d < n; d++) {
            const ArmonParams& q = b->p[d];
            const int64_t o = b->offset[d];
            launch_cfl_real<double>(b->d_states + d, b->arr[ARMON_FIELD_UMAT] + o, b->arr[ARMON_FIELD_VMAT] + o,
                                    b->arr[ARMON_FIELD_CMAT] + o, q.nx, q.ny, q.row_length, q.index_start,
                                    q.domain_size[0] / (double) q.nx, q.domain_size[1] / (double) q.ny, b->stream);
        }
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        b->cycles_enqueued = 0;
        // both graphs now, so that no later advance pays for a capture
        if (b->common.graph_chunk > 0)
            for (int parity = 0; parity < 2; parity++)
                if (!b->graph[parity]) build_graph(b, parity);
        b->begun = true;
        b->finished = false;
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_begin", e);
        return 1;
    }
}

ARMON_EXPORT int armon_batch_advance(ArmonBatch* b, int n)
{
    try {
        if (!b->begun || b->finished) throw invalid_argument("armon_batch_begin must come first");
        advance(b, n);
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_advance", e);
        return 1;
    }
}

ARMON_EXPORT int armon_batch_active(ArmonBatch* b)
{
    try {
        int k = 0;
        for (const DeviceState& s : read_states(b)) k += s.active && !s.error;
        return k;
    } catch (const std::exception& e) {
        fail("armon_batch_active", e);
        return -1;
    }
}

ARMON_EXPORT int armon_batch_run(ArmonBatch* b)
{
    try {
        if (!b->begun || b->finished) throw invalid_argument("armon_batch_begin must come first");
        const int chunk = b->common.graph_chunk > 0 ? b->common.graph_chunk : 16;
        while (true) {
            advance(b, chunk);
            bool any = false;
            for (const DeviceState& s : read_states(b)) any = any || (s.active && !s.error);
            if (!any) break;
        }
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_run", e);
        return 1;
    }
}

// Per domain, as the single-domain solver's finish: the final state goes to rho/umat/vmat/Emat,
// and pmat/cmat/gmat get the EOS of the last sweep's input state.
ARMON_EXPORT int armon_batch_finish(ArmonBatch* b)
{
    try {
        if (!b->begun) throw invalid_argument("armon_batch_begin must come first");
        if (b->finished) return 0;
        b->final_states = read_states(b);
        int status = 0;
        const int eos = b->var.eos;
        for (int d = 0; d < (int) b->p.size(); d++) {
            const DeviceState& h = b->final_states[d];
            ArmonParams& q = b->p[d];
            q.cycle = h.cycle;
            q.time = h.time;
            q.current_cycle_dt = h.current_dt;
            q.next_cycle_dt = h.next_dt;
            if (h.error) status = 1;
            const int64_t sweeps = (int64_t) h.cycle * b->sweeps_per_cycle;
            if (sweeps == 0) continue;
            const int cur = (int) (sweeps & 1);
            const int64_t o = b->offset[d];
            double* P = b->arr[ARMON_FIELD_PMAT] + o;
            double* C = b->arr[ARMON_FIELD_CMAT] + o;
            double* G = b->arr[ARMON_FIELD_GMAT] + o;
            const int fa[4] = {ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT, ARMON_FIELD_EMAT};
            const int fb[4] = {ARMON_FIELD_WORK1, ARMON_FIELD_WORK2, ARMON_FIELD_WORK3, ARMON_FIELD_WORK4};
            if (b->fp64()) {
                const int* in = cur ? fa : fb;   // input of the last sweep
                launch_eos_real<double>(b->arr[in[0]] + o, b->arr[in[1]] + o, b->arr[in[2]] + o, b->arr[in[3]] + o, P,
                                        C, G, eos, 1.4, q.nx, q.ny, q.row_length, q.index_start, b->stream);
                if (cur)
                    for (int f = 0; f < 4; f++)
                        hip_check(hipMemcpyAsync(b->arr[fa[f]] + o, b->arr[fb[f]] + o, sizeof(double) * q.nb_cells,
                                                 hipMemcpyDeviceToDevice, b->stream),
                                  "copy final state");
            } else {
                float* const* in = cur ? &b->f32[0] : &b->f32[4];
                launch_eos_real<float>(in[0] + o, in[1] + o, in[2] + o, in[3] + o, P, C, G, eos, 1.4, q.nx, q.ny,
                                       q.row_length, q.index_start, b->stream);
                for (int f = 0; f < 4; f++)
                    launch_convert_f2d(b->f32[(cur ? 4 : 0) + f] + o, b->arr[fa[f]] + o, q.nb_cells, b->stream);
            }
        }
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        b->finished = true;
        b->initialized = false;   // a new run starts from armon_batch_init
        return status;
    } catch (const std::exception& e) {
        fail("armon_batch_finish", e);
        return 2;
    }
}

ARMON_EXPORT int armon_batch_domain_result(ArmonBatch* b, int d, ArmonResult* out)
{
    if (d < 0 || d >= (int) b->p.size() || !b->finished) return 1;
    const DeviceState& h = b->final_states[d];
    memset(out, 0, sizeof *out);
    out->status = h.error ? 1 : 0;
    out->cycles = h.cycle;
    out->time = h.time;
    out->current_dt = h.current_dt;
    out->final_dt = h.next_dt;
    return 0;
}

ARMON_EXPORT int armon_batch_read_field(ArmonBatch* b, int d, int field, double* host, int64_t n)
{
    try {
        if (d < 0 || d >= (int) b->p.size()) throw out_of_range("unknown domain");
        if (field < 0 || field >= ARMON_FIELD_COUNT) throw out_of_range("unknown field");
        if (n != b->p[d].nb_cells) throw invalid_argument("host buffer size must equal the domain's nb_cells");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(host, b->arr[field] + b->offset[d], sizeof(double) * (size_t) n, hipMemcpyDeviceToHost),
                  "read field");
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_read_field", e);
        return 1;
    }
}

ARMON_EXPORT int armon_batch_kernel_times(ArmonBatch* b, int n, ArmonKernelTimes* out)
{
    try {
        if (!b->begun || b->finished) throw invalid_argument("armon_batch_begin must come first");
        memset(out, 0, sizeof *out);
        Timer t;
        t.s = b->stream;
        for (int k = 0; k < n; k++) {
            enqueue_cycle(b, b->cycles_enqueued, &t);
            b->cycles_enqueued++;
        }
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        for (size_t k = 0; k < t.ev.size(); k++) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, t.ev[k].first, t.ev[k].second), "hipEventElapsedTime");
            out->launches[t.cls[k]]++;
            out->total_ms[t.cls[k]] += ms;
        }
        return 0;
    } catch (const std::exception& e) {
        fail("armon_batch_kernel_times", e);
        return 1;
    }
}

}  // extern "C"
