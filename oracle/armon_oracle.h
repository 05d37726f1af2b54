/*
 * Armon 2D Lagrange + remap oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C (C11 + OpenMP) restatement of the reference CPU algorithm
 * (Keluaa/Armon-Kokkos, src/kernels/ + src/armon_2D.cpp, parameters.cpp,
 * ranges.cpp). It is the checker for the HIP path and the CPU timing
 * baseline ("port"). Only tests, the smoke check and the bench's CPU-baseline
 * leg may load it; the product library never links or calls it.
 *
 * Each function cites the reference file:line it follows. Arithmetic is kept
 * in the reference's operation order so that an IEEE-strict build
 * (-ffp-contract=off, no -ffast-math) is bit-comparable with the GPU path.
 *
 * Deliberate, documented deviations from a literal transcription:
 *   - Kokkos::pow(y, 2) is evaluated as y*y (what GCC and LLVM fold it to;
 *     both are the correctly rounded square).
 *   - Kokkos::pow(y, 3|4) uses armon_cr_pow3/armon_cr_pow4: a double-double
 *     evaluation rounded once, i.e. the correctly rounded power. The tests pin
 *     it against glibc pow() (which the reference's IEEE build calls).
 *   - Kokkos::min/max use the Kokkos definitions (b < a ? b : a) and
 *     (a < b ? b : a), which fix the sign of zero and NaN handling.
 *   - conservation_vars sums in a fixed, thread-count independent order with
 *     Neumaier compensation (Kokkos' sum order is backend-defined).
 *
 * Every function is pinned by the tests (tests/test_oracle.py: the reference's known
 * answers, closed forms, exact rational evaluations, invariants, brute force), except two
 * edge semantics taken from Kokkos' definitions rather than from a test of the reference:
 * the DBL_MAX identity of its Min reducer and the min/max forms above.
 */
#ifndef ARMON_ORACLE_H
#define ARMON_ORACLE_H

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t  o_idx;   /* Idx  (indexing.h:19), fixed to 64 bits */
typedef uint64_t o_uidx;  /* UIdx (indexing.h:14) */

/* indexing.h:28-33 */
typedef struct { o_idx start, end; } ORange;
/* indexing.h:39-45 */
typedef struct { o_idx start, step; } OInnerRange1D;
/* indexing.h:60-73 */
typedef struct { o_idx main_range_start, main_range_step, row_range_start; o_uidx row_range_length; } OInnerRange2D;
/* indexing.h:93-116 (col_end/row_end inclusive) */
typedef struct { int64_t col_start, col_step, col_end, row_start, row_step, row_end; } ODomainRange;

/* Enumerations, same integer values as the reference. */
enum { O_LIM_NONE = 0, O_LIM_MINMOD = 1, O_LIM_SUPERBEE = 2 };            /* limiters.h:8-12 */
enum { O_TEST_SOD = 0, O_TEST_SOD_Y = 1, O_TEST_SOD_CIRC = 2,
       O_TEST_BIZARRIUM = 3, O_TEST_SEDOV = 4 };                          /* test_cases.h:8-14 */
enum { O_SCHEME_GODUNOV = 0, O_SCHEME_GAD = 1 };                          /* parameters.h:17-19 */
enum { O_PROJ_NONE = 0, O_PROJ_EULER = 1, O_PROJ_EULER_2ND = 2 };         /* parameters.h:25-27 */
enum { O_SPLIT_SEQUENTIAL = 0, O_SPLIT_SEQUENTIAL_SYM = 1, O_SPLIT_STRANG = 2,
       O_SPLIT_X_ONLY = 3, O_SPLIT_Y_ONLY = 4 };                          /* parameters.h:29-31 */
enum { O_AXIS_X = 0, O_AXIS_Y = 1 };                                      /* common.h:30-33 */
enum { O_SIDE_LEFT = 0, O_SIDE_RIGHT = 1, O_SIDE_TOP = 2, O_SIDE_BOTTOM = 3 }; /* common.h:36-41 */

/* ---- index helpers (indexing.h / indexing.cpp) ---- */
o_idx o_scale_index_1d(const OInnerRange1D* r, o_uidx i);
o_idx o_scale_index_2d(const OInnerRange2D* r, o_uidx i);
uint64_t o_dr_col_length(const ODomainRange* d);
uint64_t o_dr_row_length(const ODomainRange* d);
uint64_t o_dr_length(const ODomainRange* d);
int64_t  o_dr_begin(const ODomainRange* d);
int64_t  o_dr_end(const ODomainRange* d);
void o_dr_inflate_dir(ODomainRange* d, int axis, int64_t n);
void o_dr_expand_dir(ODomainRange* d, int axis, int64_t n);
void o_dr_iter1d(const ODomainRange* d, ORange* r, OInnerRange1D* ir);
void o_dr_direct_iter1d(const ODomainRange* d, ORange* r, OInnerRange1D* ir);
void o_dr_iter2d(const ODomainRange* d, ORange* r, OInnerRange2D* ir);

/* ---- correctly rounded small integer powers (see header comment) ---- */
double armon_cr_pow3(double y);
double armon_cr_pow4(double y);

double o_limiter(int limiter, double r);   /* limiters.h:15-34 */

/* ---- the 14 kernels (src/kernels/kernels.h:10-86), on flat host arrays ---- */
void o_acoustic(const ORange* range, const OInnerRange2D* ir, o_idx s,
                double* ustar, double* pstar,
                const double* rho, const double* u, const double* pmat, const double* cmat);
int  o_acoustic_GAD(const ORange* range, const OInnerRange2D* ir, o_idx s, double dt, double dx,
                    double* ustar, double* pstar,
                    const double* rho, const double* u, const double* pmat, const double* cmat,
                    int limiter);
void o_perfect_gas_EOS(const ORange* range, const OInnerRange2D* ir, double gamma,
                       const double* rho, const double* Emat, const double* umat, const double* vmat,
                       double* pmat, double* cmat, double* gmat);
void o_bizarrium_EOS(const ORange* range, const OInnerRange2D* ir,
                     const double* rho, const double* umat, const double* vmat, const double* Emat,
                     double* pmat, double* cmat, double* gmat);
void o_cell_update(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                   const double* ustar, const double* pstar, double* rho, double* u, double* Emat);
void o_euler_projection(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                        const double* ustar, double* rho, double* umat, double* vmat, double* Emat,
                        const double* adv_rho, const double* adv_urho,
                        const double* adv_vrho, const double* adv_Erho);
void o_advection_first_order(const ORange* range, const OInnerRange2D* ir, o_idx s, double dt,
                             const double* ustar, const double* rho, const double* umat,
                             const double* vmat, const double* Emat,
                             double* adv_rho, double* adv_urho, double* adv_vrho, double* adv_Erho);
void o_advection_second_order(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                              const double* ustar, const double* rho, const double* umat,
                              const double* vmat, const double* Emat,
                              double* adv_rho, double* adv_urho, double* adv_vrho, double* adv_Erho);
void o_boundary_conditions(const ORange* range, const OInnerRange1D* ir, o_idx stencil_width, o_idx disp,
                           double u_factor, double v_factor,
                           double* rho, double* umat, double* vmat, double* pmat,
                           double* cmat, double* gmat, double* Emat);
void o_read_border_array(const ORange* range, const OInnerRange2D* ir, o_idx side_length, o_idx nghost,
                         const double* rho, const double* umat, const double* vmat, const double* pmat,
                         const double* cmat, const double* gmat, const double* Emat,
                         double* value_array);
void o_write_border_array(const ORange* range, const OInnerRange2D* ir, o_idx side_length, o_idx nghost,
                          double* rho, double* umat, double* vmat, double* pmat,
                          double* cmat, double* gmat, double* Emat,
                          const double* value_array);
/* High/low region states of a test (init_test.cpp:12-48). */
typedef struct { double high_rho, low_rho, high_E, low_E, high_u, low_u, high_v, low_v; } OTestState;
void o_reference_test_state(int test, double test_option, OTestState* out);
int  o_init_test_states(const ORange* range, const OInnerRange2D* ir,
                        o_idx row_length, o_idx nb_ghosts, o_idx nx, o_idx ny, double sx, double sy,
                        double ox, double oy, o_idx pos_x, o_idx pos_y, o_idx g_nx, o_idx g_ny,
                        double* x, double* y, double* rho, double* Emat, double* umat, double* vmat,
                        double* domain_mask, double* pmat, double* cmat, double* gmat,
                        double* ustar, double* pstar, int test, bool debug_indexes, double test_option,
                        const OTestState* states);
int  o_init_test(const ORange* range, const OInnerRange2D* ir,
                 o_idx row_length, o_idx nb_ghosts, o_idx nx, o_idx ny, double sx, double sy,
                 double ox, double oy, o_idx pos_x, o_idx pos_y, o_idx g_nx, o_idx g_ny,
                 double* x, double* y, double* rho, double* Emat, double* umat, double* vmat,
                 double* domain_mask, double* pmat, double* cmat, double* gmat,
                 double* ustar, double* pstar, int test, bool debug_indexes, double test_option);
double o_dt_CFL(const ORange* range, const OInnerRange1D* ir, double dx, double dy,
                const double* umat, const double* vmat, const double* cmat, const double* domain_mask);
void o_conservation_vars(const ORange* range, const OInnerRange1D* ir, double dx,
                         const double* rho, const double* Emat, const double* domain_mask,
                         double* total_mass, double* total_energy);

/* ---- solver parameters (parameters.h:34-108) ---- */
typedef struct {
    int test, scheme, riemann, projection, limiter, current_axis, axis_splitting;
    int nb_ghosts, nx, ny;
    double dx, cfl, Dt;
    int stencil_width;            /* -1: min_ghosts() */
    int cst_dt, dt_on_even_cycles;
    double domain_size[2];        /* <= 0: test default */
    double domain_origin[2];      /* NaN: test default */
    /* indexing (derived) */
    int row_length, col_length, nb_cells, ideb, ifin, index_start, idx_row, idx_col;
    int s;
    /* solver state */
    int cycle;
    double cycle_dt, current_cycle_dt, next_cycle_dt, time;
    /* bounds */
    int max_cycles;
    double max_time;
    int verbose;
    /* boundary velocity factors (u, v) per side (O_SIDE_*), used instead of the test's
     * (test_cases.h) when bc_override is set: a configuration outside the reference's tests,
     * such as reflective walls around a Sedov blast at a corner */
    int bc_override;
    double bc_factors[4][2];
} OParams;

void   o_params_defaults(OParams* p);                 /* parameters.h:34-97 member initialisers */
int    o_min_ghosts(const OParams* p);                /* parameters.cpp:11-18 */
void   o_params_init(OParams* p);                     /* parameters.cpp:21-34 */
int    o_params_check(const OParams* p);              /* parameters.cpp:70-109 (1 = ok) */
void   o_update_axis(OParams* p, int axis);           /* parameters.cpp:195-210 */
int    o_split_axes(const OParams* p, int axes[3], double factors[3]); /* parameters.cpp:212-255 */
void   o_test_bc(int test, int side, double* u_factor, double* v_factor); /* test_cases.h:25-121 */
double o_test_default_cfl(int test);
double o_test_default_max_time(int test);

/* ---- iteration domains (ranges.cpp) ---- */
int64_t      o_index_1D(const OParams* p, int i, int j);          /* ranges.h:44-47 */
ODomainRange o_real_domain(const OParams* p);                     /* ranges.cpp:5-11 */
ODomainRange o_domain_fluxes(const OParams* p);                   /* ranges.cpp:14-30 */
ODomainRange o_domain_cell_update(const OParams* p);              /* ranges.cpp:33-46 */
ODomainRange o_domain_advection(const OParams* p);                /* ranges.cpp:49-54 */
ODomainRange o_complete_domain(const OParams* p);                 /* ranges.cpp:57-63 */
ODomainRange o_boundary_conditions_domain(const OParams* p, int side, int* disp); /* ranges.cpp:66-106 */

/* ---- solver data (data.h:15-39): 16 arrays of nb_cells ---- */
typedef struct {
    double *x, *y, *rho, *umat, *vmat, *Emat, *pmat, *cmat, *gmat, *ustar, *pstar;
    double *work_array_1, *work_array_2, *work_array_3, *work_array_4;
    double *domain_mask;
} OData;

int  o_data_alloc(OData* d, int64_t nb_cells);        /* 64-byte aligned */
void o_data_free(OData* d);

/* ---- driver (armon_2D.cpp) ---- */
double o_test_option(const OParams* p);                                  /* armon_2D.cpp:116-118 */
void   o_init_test_params(const OParams* p, OData* d, bool debug_indexes); /* armon_2D.cpp:112-127 */
void   o_init_test_params_states(const OParams* p, OData* d, const OTestState* states);
void   o_update_EOS(const OParams* p, OData* d);                          /* armon_2D.cpp:85-101 */
void   o_boundary_conditions_all(const OParams* p, OData* d);             /* armon_2D.cpp:130-151 */
void   o_numerical_fluxes(const OParams* p, OData* d);                    /* armon_2D.cpp:70-82 */
void   o_cell_update_params(const OParams* p, OData* d);                  /* armon_2D.cpp:104-109 */
void   o_projection_remap(const OParams* p, OData* d);                    /* armon_2D.cpp:188-209 */
double o_local_time_step(const OParams* p, OData* d, double prev_dt);     /* armon_2D.cpp:212-226 */
int    o_time_step(OParams* p, OData* d);                                 /* armon_2D.cpp:229-252 (0 = ok) */
int    o_solver_cycle(OParams* p, OData* d);                              /* armon_2D.cpp:304-324 */
void   o_conservation_vars_params(const OParams* p, OData* d, double* mass, double* energy); /* :287-298 */
/* armon_2D.cpp:327-386. Returns 0 on success, 1 on an invalid time step.
 * loop_time is the wall time of the loop in seconds. */
int    o_time_loop(OParams* p, OData* d, double* loop_time, double* final_dt, int* cycles);

/* OpenMP helpers for the CPU baseline */
int  o_num_threads(void);
void o_set_num_threads(int n);
/* 1 if this build uses IEEE-strict arithmetic (the checker), 0 for the fast-math baseline build */
int  o_is_strict(void);

#ifdef __cplusplus
}
#endif

#endif /* ARMON_ORACLE_H */
