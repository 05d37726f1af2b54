/*
 * Armon MI355X library — extensions beside the drop-in kernel C-ABI (kernels.h).
 *
 *  1. Plumbing: library stream, device memory and copies (the caller owns memory
 *     passed to the kernels; these helpers exist so that C/ctypes callers need no
 *     other GPU runtime binding).
 *  2. dt_CFL_device: the CFL reduction with its result left in device memory.
 *  3. The host driver: Params, iteration domains and the time loop of the
 *     reference (src/parameters.cpp, src/ranges.cpp, src/armon_2D.cpp), with two
 *     engines:
 *       ARMON_ENGINE_STAGED  one launch per stage through the drop-in entry points,
 *                            host-side dt (the reference's schedule);
 *       ARMON_ENGINE_FUSED   one fused Lagrange+remap sweep kernel per axis, the
 *                            CFL reduction folded into the last sweep of a cycle,
 *                            dt/time/cycle kept on the device, cycles replayed from
 *                            a captured hipGraph.
 *     The fused engine runs in FP64 (bit-compatible with the staged engine) or in
 *     FP32-mixed precision (FP32 storage and flux arithmetic; FP64 EOS, dt and
 *     conservation sums).
 *
 * All functions return 0 on success unless stated otherwise; on failure the error
 * also goes through raise_exception_handler and armon_last_error().
 */
#ifndef ARMON_HIP_ARMON_H
#define ARMON_HIP_ARMON_H

#include <stddef.h>

#include "armon/kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- plumbing ---------------------------------------------------------- */

/* Stream used by every launch of the library (a hipStream_t; NULL = default stream). */
ARMON_EXPORT int armon_set_stream(void* hip_stream);
ARMON_EXPORT void* armon_get_stream(void);

ARMON_EXPORT int armon_device_count(void);
ARMON_EXPORT void* armon_malloc(size_t bytes);            /* device memory, 256-byte aligned; NULL on failure */
ARMON_EXPORT int armon_free(void* device_ptr);
ARMON_EXPORT int armon_memcpy_htod(void* dst, const void* src, size_t bytes);  /* synchronous */
ARMON_EXPORT int armon_memcpy_dtoh(void* dst, const void* src, size_t bytes);  /* synchronous */
ARMON_EXPORT int armon_memcpy_dtod(void* dst, const void* src, size_t bytes);  /* async on the library stream */
ARMON_EXPORT int armon_synchronize(void);                 /* waits for the library stream */
ARMON_EXPORT const char* armon_last_error(void);          /* "" when no error was recorded */
ARMON_EXPORT void armon_clear_error(void);
ARMON_EXPORT const char* armon_version(void);

/* ---- device-resident CFL reduction -------------------------------------- */

/* Same reduction as dt_CFL (reductions.cpp:8-24) but asynchronous: the minimum is
 * written to *dt_device (a device address) on the library stream. */
ARMON_EXPORT void dt_CFL_device(ARMON_CREF(Range) range, ARMON_CREF(InnerRange1D) inner_range, flt_t dx, flt_t dy,
                                ARMON_CREF(view) umat, ARMON_CREF(view) vmat, ARMON_CREF(view) cmat,
                                ARMON_CREF(mask_view) domain_mask, double* dt_device);

/* ---- allocation-free conservation sums ------------------------------------ */

/* conservation_vars (same arguments, results and error behaviour) that allocates nothing at
 * any call, for callers that cannot allocate device memory (tight loops, harnesses that forbid
 * hipMalloc). It has no counterpart in the reference. Its partial sums live in a static device
 * array of the library, 4 x 1024 + 2 doubles (32,784 bytes) per device, held until the process
 * ends; any range fits, since the sum folds any number of cells into at most 1024 block
 * partials. Calls that overlap on one device share that array and race. */
ARMON_EXPORT void conservation_vars_noalloc(ARMON_CREF(Range) range, ARMON_CREF(InnerRange1D) inner_range, flt_t dx,
                                            ARMON_CREF(view) rho, ARMON_CREF(view) Emat,
                                            ARMON_CREF(mask_view) domain_mask,
                                            ARMON_REF(flt_t) total_mass, ARMON_REF(flt_t) total_energy);

/* ---- initial states -------------------------------------------------------- */

/* High/low region states of a test case (init_test.cpp:12-48, TestParams of test_cases.h:17-22). */
typedef struct ArmonTestState {
    double high_rho, low_rho, high_E, low_E, high_u, low_u, high_v, low_v;
} ArmonTestState;

/* The reference states of `test` (test_option: the Sedov radius, test_cases.h:118-120). */
ARMON_EXPORT void armon_reference_test_state(int test, double test_option, ArmonTestState* out);

/* init_test with caller-given high/low states instead of the test's reference ones (used by
 * the batched ensemble, whose domains perturb their initial states). Same regions, masks and
 * initial p, c, g, ustar, pstar as init_test; no debug-index mode. */
ARMON_EXPORT void init_test_states(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range,
                                   Idx row_length, Idx nb_ghosts, Idx nx, Idx ny, flt_t sx,
                                   flt_t sy, flt_t ox, flt_t oy, Idx pos_x, Idx pos_y, Idx g_nx, Idx g_ny,
                                   ARMON_REF(view) x, ARMON_REF(view) y, ARMON_REF(view) rho, ARMON_REF(view) Emat,
                                   ARMON_REF(view) umat, ARMON_REF(view) vmat, ARMON_REF(mask_view) domain_mask,
                                   ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat,
                                   ARMON_REF(view) ustar, ARMON_REF(view) pstar,
                                   Test test, flt_t test_option, ARMON_CREF(ArmonTestState) states);

/* ---- measurement --------------------------------------------------------- */

/* Bandwidth (GB/s, bytes read + written per second) of a plain device copy of `bytes`
 * bytes repeated `iters` times: the achievable HBM rate the sweeps are compared with.
 * Returns a negative value on error. */
ARMON_EXPORT double armon_copy_bandwidth(size_t bytes, int iters);

/* ---- parameters (src/parameters.h:17-108) -------------------------------- */

enum { ARMON_SCHEME_GODUNOV = 0, ARMON_SCHEME_GAD = 1 };
enum { ARMON_RIEMANN_ACOUSTIC = 0 };
enum { ARMON_PROJ_NONE = 0, ARMON_PROJ_EULER = 1, ARMON_PROJ_EULER_2ND = 2 };
enum { ARMON_SPLIT_SEQUENTIAL = 0, ARMON_SPLIT_SEQUENTIAL_SYM = 1, ARMON_SPLIT_STRANG = 2,
       ARMON_SPLIT_X_ONLY = 3, ARMON_SPLIT_Y_ONLY = 4 };
enum { ARMON_AXIS_X = 0, ARMON_AXIS_Y = 1 };
enum { ARMON_SIDE_LEFT = 0, ARMON_SIDE_RIGHT = 1, ARMON_SIDE_TOP = 2, ARMON_SIDE_BOTTOM = 3 };
enum { ARMON_ENGINE_STAGED = 0, ARMON_ENGINE_FUSED = 1 };
enum { ARMON_PREC_FP64 = 0, ARMON_PREC_FP32_MIXED = 1 };

typedef struct ArmonParams {
    int test, scheme, riemann, projection, limiter, current_axis, axis_splitting;
    int nb_ghosts, nx, ny;
    double dx, cfl, Dt;
    int stencil_width;              /* -1: min_ghosts() */
    int cst_dt, dt_on_even_cycles;
    double domain_size[2];          /* <= 0: test default */
    double domain_origin[2];        /* NaN: test default */
    /* indexing (set by armon_params_init) */
    int row_length, col_length, nb_cells, ideb, ifin, index_start, idx_row, idx_col;
    int s;
    /* solver state */
    int cycle;
    double cycle_dt, current_cycle_dt, next_cycle_dt, time;
    /* bounds */
    int max_cycles;
    double max_time;
    /* output */
    int verbose;
    int write_output, write_ghosts, write_throughput, output_precision;
    const char* output_file;
    /* execution options of this implementation */
    int engine;                     /* ARMON_ENGINE_* */
    int precision;                  /* ARMON_PREC_* (fused engine only) */
    int graph_chunk;                /* cycles per captured graph launch (fused engine); 0: no graph.
                                       Advancing n cycles launches n / graph_chunk graphs (one graph
                                       per parity of the start cycle, captured at begin), then the
                                       remaining n % graph_chunk cycles launch by launch. */
    /* per-step comparison with reference files (parameters.h:95-97; staged engine only) */
    int compare;
    double comparison_tolerance;    /* relative, default 1e-13 */
    int y_run_rows;                 /* fused Y sweep: rows per marching run; 0: chosen from the domain size */
    /* boundary velocity factors (u, v) per side (ARMON_SIDE_*), used instead of the test's
     * (test_cases.h:36-38 and the per-test overrides) when bc_override is set: a configuration
     * outside the reference's tests, such as reflective walls around a Sedov blast at a corner.
     * Both engines and the batched ensemble's common parameters honour it. */
    int bc_override;
    double bc_factors[4][2];
} ArmonParams;

ARMON_EXPORT void armon_params_default(ArmonParams* p);              /* parameters.h:34-97 defaults */
/* parameters.cpp:21-67. Returns 1 (and leaves the indexing fields unset) for an unknown test or
   a domain of 2^31 cells or more, ghosts included, whose int indexing would wrap around. */
ARMON_EXPORT int armon_params_init(ArmonParams* p);
ARMON_EXPORT int armon_params_check(const ArmonParams* p);           /* parameters.cpp:70-109: 1 if valid */
ARMON_EXPORT int armon_min_ghosts(const ArmonParams* p);             /* parameters.cpp:11-18 */
ARMON_EXPORT void armon_update_axis(ArmonParams* p, int axis);       /* parameters.cpp:195-210 */
ARMON_EXPORT int armon_split_axes(const ArmonParams* p, int axes[3], double factors[3]); /* parameters.cpp:212-255 */
ARMON_EXPORT void armon_test_boundary_condition(int test, int side, double* u_factor, double* v_factor);
ARMON_EXPORT double armon_test_option(const ArmonParams* p);         /* test_cases.h:32, 118-120 */

/* ---- iteration domains (src/kernels/indexing.h:93-116, src/ranges.cpp) ---- */

typedef struct DomainRange {
    int64_t col_start, col_step, col_end;   /* *_end inclusive */
    int64_t row_start, row_step, row_end;
} DomainRange;

ARMON_EXPORT uint64_t armon_domain_col_length(const DomainRange* d);
ARMON_EXPORT uint64_t armon_domain_row_length(const DomainRange* d);
ARMON_EXPORT uint64_t armon_domain_length(const DomainRange* d);
ARMON_EXPORT int64_t armon_domain_begin(const DomainRange* d);
ARMON_EXPORT int64_t armon_domain_end(const DomainRange* d);
ARMON_EXPORT void armon_domain_iter1D(const DomainRange* d, Range* r, InnerRange1D* ir);
ARMON_EXPORT void armon_domain_direct_iter1D(const DomainRange* d, Range* r, InnerRange1D* ir);
ARMON_EXPORT void armon_domain_iter2D(const DomainRange* d, Range* r, InnerRange2D* ir);

ARMON_EXPORT int64_t armon_index_1D(const ArmonParams* p, int i, int j);
ARMON_EXPORT DomainRange armon_real_domain(const ArmonParams* p);
ARMON_EXPORT DomainRange armon_domain_fluxes(const ArmonParams* p);
ARMON_EXPORT DomainRange armon_domain_cell_update(const ArmonParams* p);
ARMON_EXPORT DomainRange armon_domain_advection(const ArmonParams* p);
ARMON_EXPORT DomainRange armon_complete_domain(const ArmonParams* p);
ARMON_EXPORT DomainRange armon_boundary_conditions_domain(const ArmonParams* p, int side, int* disp);

/* ---- solver (src/armon_2D.cpp) ------------------------------------------ */

/* Field identifiers, in the order of data.h:103-106 plus the mask. */
enum { ARMON_FIELD_X = 0, ARMON_FIELD_Y, ARMON_FIELD_RHO, ARMON_FIELD_UMAT, ARMON_FIELD_VMAT,
       ARMON_FIELD_EMAT, ARMON_FIELD_PMAT, ARMON_FIELD_CMAT, ARMON_FIELD_GMAT, ARMON_FIELD_USTAR,
       ARMON_FIELD_PSTAR, ARMON_FIELD_WORK1, ARMON_FIELD_WORK2, ARMON_FIELD_WORK3, ARMON_FIELD_WORK4,
       ARMON_FIELD_DOMAIN_MASK, ARMON_FIELD_COUNT };

typedef struct ArmonResult {
    int status;          /* 0 ok, 1 invalid time step (the reference exits), 2 other error,
                            3 the per-step comparison found a difference (the loop stopped there) */
    int cycles;
    double time;
    double current_dt;   /* dt of the last cycle */
    double final_dt;     /* next_cycle_dt, as printed by the reference */
    double loop_time;    /* seconds, EOS_init + loop + final fence (armon_2D.cpp:334-373) */
    double grind_time;   /* microseconds per cell per cycle */
    double initial_mass, initial_energy, final_mass, final_energy;
} ArmonResult;

typedef struct ArmonKernelTimes {
    int launches[4];      /* [0] X sweeps, [1] Y sweeps, [2] time-step kernels, [3] other */
    double total_ms[4];   /* summed HIP-event durations per kernel class */
} ArmonKernelTimes;

typedef struct ArmonSolver ArmonSolver;

/* Allocates the device data for `p` (which must have been through armon_params_init). */
ARMON_EXPORT ArmonSolver* armon_solver_create(const ArmonParams* p);
ARMON_EXPORT void armon_solver_destroy(ArmonSolver* s);
ARMON_EXPORT int armon_solver_get_params(const ArmonSolver* s, ArmonParams* out);
/* armon_2D.cpp:112-127 */
ARMON_EXPORT int armon_solver_init_test(ArmonSolver* s, bool debug_indexes);
/* armon_2D.cpp:327-386: EOS_init, then cycles while time < max_time && cycle < max_cycles. */
ARMON_EXPORT int armon_solver_time_loop(ArmonSolver* s, ArmonResult* out);
/* Enqueues up to n more cycles (still bounded by max_time/max_cycles on the device)
 * without synchronising; armon_solver_time_loop or armon_solver_begin must come first. */
ARMON_EXPORT int armon_solver_begin(ArmonSolver* s);
ARMON_EXPORT int armon_solver_advance(ArmonSolver* s, int n);
ARMON_EXPORT int armon_solver_finish(ArmonSolver* s, ArmonResult* out);
/* Runs n cycles launch by launch, timing every kernel with HIP events on the solver stream. */
ARMON_EXPORT int armon_solver_kernel_times(ArmonSolver* s, int n, ArmonKernelTimes* out);
/* Per-stage device time since the solver was created, under the reference's stage labels
 * (armon_2D.cpp:57-67, 306-322: init_test, dt_CFL, EOS, boundary_conditions, numerical_fluxes,
 * cell_update, euler_proj for the staged engine; the graph launches and eagerly launched cycles
 * of the fused engine), each stage bracketed by HIP events on the solver stream. Recorded when
 * the parameters' verbose < 3, when armon_run prints the table as the reference does. Stages are
 * numbered 0 .. armon_solver_stage_count() - 1 in label order; label strings live as long as the
 * solver. */
ARMON_EXPORT int armon_solver_stage_count(ArmonSolver* s);
ARMON_EXPORT int armon_solver_stage_time(ArmonSolver* s, int k, const char** label, double* ms, int64_t* calls);
/* Copies a whole field (nb_cells values) to/from the host as doubles. For the fused
 * engine, fields are the current state; PMAT/CMAT hold the EOS of the input of the
 * last sweep, exactly as in the reference. The fused engine keeps the real cells only: its
 * sweeps make the ghost cells in registers and never store them, so ghost cells read back
 * hold what init_test (or the last write_field) left there; the staged engine's ghost cells
 * are the reference's. */
ARMON_EXPORT int armon_solver_read_field(ArmonSolver* s, int field, double* host, int64_t n);
ARMON_EXPORT int armon_solver_write_field(ArmonSolver* s, int field, const double* host, int64_t n);
ARMON_EXPORT int armon_solver_conservation(ArmonSolver* s, double* mass, double* energy);
ARMON_EXPORT void* armon_solver_stream(ArmonSolver* s);

/* armon_2D.cpp:389-431: allocate, init, loop, optional output. */
ARMON_EXPORT int armon_run(ArmonParams* p, ArmonResult* out);

/* ---- batched ensemble ------------------------------------------------------
 * Many independent domains of different sizes advanced together: every sweep of a cycle
 * is ONE persistent launch over the work items (row tiles for X, column runs for Y) of all
 * domains, each domain with its own dt, time, cycle count and end time, all on the device;
 * cycles are replayed from a captured hipGraph. Each domain follows exactly the reference
 * time loop (armon_2D.cpp:327-386) on its own parameters, so a domain's result is the
 * single-domain result. This has no counterpart in the reference, which runs one domain
 * per process. */

typedef struct ArmonDomainSpec {
    int test;                 /* Sod, Sod_y, Sod_circ or Sedov (perfect gas); or Bizarrium for every domain */
    int nx, ny;               /* real cells */
    double max_time;          /* end time; <= 0: the test's default */
    ArmonTestState state;     /* high/low region states (armon_reference_test_state for the reference ones) */
} ArmonDomainSpec;

/* `n` domains drawn from splitmix64(seed): area log-uniform in [min_area, max_area] cells,
 * aspect ratio log-uniform in [1, max_aspect] with a random orientation, test uniform over
 * {Sod, Sod_circ, Sedov}, and the high/low densities and pressures of the reference states
 * each multiplied by a factor uniform in [0.8, 1.2] (E scaled so that p follows). Sides are
 * at least 16 cells. Returns 0, or 1 on invalid arguments. */
ARMON_EXPORT int armon_batch_random_specs(uint64_t seed, int n, int64_t min_area, int64_t max_area,
                                          double max_aspect, ArmonDomainSpec* out);

/* Launch configuration of the batched sweeps, chosen by armon_batch_autotune: the tile shape
 * and the cells each thread computes (an X tile is tile_x cells along a row, computed by
 * tile_x / x_cells_per_thread threads; a Y item is a wave of 64 columns, each thread marching
 * y_run_rows cells down its column) and the launch grids. */
typedef struct ArmonBatchTuning {
    int y_run_rows;           /* rows a wave marches per Y work item (64 columns wide) */
    int grid_x, grid_y;       /* blocks of the X and Y launches (persistent when below the item count) */
    int items_x, items_y;     /* work items per X / Y sweep (outputs) */
    double ms_x, ms_y;        /* measured time per X / Y sweep of this configuration (autotune output) */
    int tile_x;               /* cells per X tile: 128, 256 or 512 (0 in set_tuning: 256) */
    int x_cells_per_thread;   /* cells per thread of the X tile: 1, or 2 for tiles of 256 or 512 cells
                                 with one block per tile (grid_x <= 0); 0 in set_tuning: 1 */
} ArmonBatchTuning;

typedef struct ArmonBatch ArmonBatch;

/* `common` (armon_params_default plus overrides) gives the scheme, limiter, projection,
 * splitting, ghosts, stencil width, Dt/cst_dt, dt_on_even_cycles, max_cycles, precision and
 * graph_chunk of every domain, and its cfl when positive (else each test's default);
 * `specs` the test, size, states and end time of each domain. */
ARMON_EXPORT ArmonBatch* armon_batch_create(const ArmonParams* common, int n, const ArmonDomainSpec* specs);
ARMON_EXPORT void armon_batch_destroy(ArmonBatch* b);
ARMON_EXPORT int armon_batch_count(const ArmonBatch* b);
ARMON_EXPORT int64_t armon_batch_cells(const ArmonBatch* b);            /* real cells of all domains */
ARMON_EXPORT int armon_batch_domain_params(const ArmonBatch* b, int d, ArmonParams* out);
ARMON_EXPORT void* armon_batch_stream(ArmonBatch* b);
/* init_test_states of every domain (armon_2D.cpp:112-127). */
ARMON_EXPORT int armon_batch_init(ArmonBatch* b);
/* Times the sweeps over the initial state for the candidate configurations (X tile width, Y
 * run length, persistent or one-block-per-item grids) and keeps the fastest. After init, before
 * begin. */
ARMON_EXPORT int armon_batch_autotune(ArmonBatch* b, ArmonBatchTuning* out);
/* Sets the configuration by hand (grid_x/grid_y <= 0: one block per item); before begin. */
ARMON_EXPORT int armon_batch_set_tuning(ArmonBatch* b, const ArmonBatchTuning* t);
ARMON_EXPORT int armon_batch_get_tuning(const ArmonBatch* b, ArmonBatchTuning* out);
ARMON_EXPORT int armon_batch_begin(ArmonBatch* b);
/* Enqueues up to n more cycles of every still-running domain, without synchronising. */
ARMON_EXPORT int armon_batch_advance(ArmonBatch* b, int n);
/* Advances until every domain has reached its end time or max_cycles (or failed). */
ARMON_EXPORT int armon_batch_run(ArmonBatch* b);
/* Number of domains still running (synchronises). */
ARMON_EXPORT int armon_batch_active(ArmonBatch* b);
/* Puts every domain's final state into its rho/umat/vmat/Emat and the EOS of its last
 * sweep's input into pmat/cmat/gmat, as the single-domain solver does. Returns 0, or 1 if
 * some domain stopped on an invalid time step. */
ARMON_EXPORT int armon_batch_finish(ArmonBatch* b);
/* cycles, time, current_dt, final_dt and status of domain d (after finish). */
ARMON_EXPORT int armon_batch_domain_result(ArmonBatch* b, int d, ArmonResult* out);
/* Copies field `field` of domain d (its nb_cells values) to the host. */
ARMON_EXPORT int armon_batch_read_field(ArmonBatch* b, int d, int field, double* host, int64_t n);
/* Runs n cycles launch by launch, timing every kernel with HIP events. After begin. */
ARMON_EXPORT int armon_batch_kernel_times(ArmonBatch* b, int n, ArmonKernelTimes* out);

#ifdef __cplusplus
}
#endif

#endif /* ARMON_HIP_ARMON_H */
