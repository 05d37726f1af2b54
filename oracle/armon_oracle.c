/*
 * Armon 2D Lagrange + remap oracle — TEST INFRASTRUCTURE ONLY.
 * See armon_oracle.h for scope and the list of deliberate deviations.
 * Citations are to the reference repository (Keluaa/Armon-Kokkos).
 */
#define _GNU_SOURCE
#include "armon_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PARALLEL_FOR _Pragma("omp parallel for schedule(static)")

/* Iteration of a Range scaled through an InnerRange2D, declaring the array index `i` for the
 * loop body that follows. The checker (liboracle.so) walks the linear range as the reference's
 * default build does (iter_lin, indexing.h:120-125), one division per element
 * (o_scale_index_2d). The throughput build (liboracle_fast.so, -DORACLE_2D_ITER=1) walks it as
 * the reference's USE_2D_ITER build does (iter_2d, indexing.h:140-155): the rows of the inner
 * range are shared among the threads and each row's cells run in a vector loop, with no
 * division per element. Every element is computed the same way in both; only the order
 * differs, which no kernel below depends on. The loop body must not modify the index. */
#if ORACLE_2D_ITER
#define FOR_2D(range, r, i)                                                                             \
    const o_idx len_ = (o_idx) (r).row_range_length;                                                    \
    const o_idx k0_ = (range)->start, k1_ = (range)->end;                                               \
    const o_idx row0_ = (len_ > 0 && k1_ > k0_) ? k0_ / len_ : 0;                                       \
    const o_idx row1_ = (len_ > 0 && k1_ > k0_) ? (k1_ - 1) / len_ : -1;                                \
    const o_idx base_ = (r).row_range_start + (r).main_range_start;                                     \
    PARALLEL_FOR                                                                                        \
    for (o_idx row_ = row0_; row_ <= row1_; row_++)                                                     \
        _Pragma("omp simd")                                                                             \
        for (o_idx i = base_ + row_ * (r).main_range_step + (row_ == row0_ ? k0_ - row0_ * len_ : 0);    \
             i < base_ + row_ * (r).main_range_step + (row_ == row1_ ? k1_ - row1_ * len_ : len_); i++)
#elif ORACLE_MD_ITER
/* The throughput build's other iteration, as the reference's USE_MD_ITER + BALANCE_MD_ITER build
 * walks it (iter_md, indexing.h:157-189): the bounding rectangle of the range in (row, column)
 * coordinates of the main step, its rows grouped in tiles of o_md_tile_rows() rows (the
 * balanced split of indexing.cpp:113-191) shared statically among the threads. The column tile,
 * left to Kokkos' default by the reference, is the whole row here, a vector loop. */
static o_idx o_md_tile_rows(o_idx rows);
#define FOR_2D(range, r, i)                                                                             \
    const int nonempty_ = (range)->end > (range)->start;                                                \
    const o_idx first_ = nonempty_ ? o_scale_index_2d(&(r), (o_uidx) (range)->start) : 0;               \
    const o_idx last_ = nonempty_ ? o_scale_index_2d(&(r), (o_uidx) ((range)->end - 1)) : 0;            \
    const o_idx step_ = (r).main_range_step;                                                            \
    const o_idx y0_ = first_ / step_, x0_ = first_ % step_;                                             \
    const o_idx y1_ = last_ / step_ + 1, x1_ = last_ % step_ + 1;                                       \
    const o_idx ty_ = nonempty_ ? o_md_tile_rows(y1_ - y0_) : 1;                                        \
    const o_idx nt_ = nonempty_ ? (y1_ - y0_ + ty_ - 1) / ty_ : 0;                                      \
    PARALLEL_FOR                                                                                        \
    for (o_idx t_ = 0; t_ < nt_; t_++)                                                                  \
        for (o_idx y_ = y0_ + t_ * ty_; y_ < y0_ + (t_ + 1) * ty_ && y_ < y1_; y_++)                    \
            _Pragma("omp simd")                                                                         \
            for (o_idx i = y_ * step_ + x0_; i < y_ * step_ + x1_; i++)
#else
#define FOR_2D(range, r, i)                                                                             \
    PARALLEL_FOR                                                                                        \
    for (o_idx k_ = (range)->start; k_ < (range)->end; k_++)                                            \
        for (o_idx i = o_scale_index_2d(&(r), (o_uidx) k_), once_ = 1; once_; once_ = 0)
#endif

/* Kokkos::min / Kokkos::max on arithmetic types: std::min / std::max semantics. */
static inline double kmin(double a, double b) { return (b < a) ? b : a; }
static inline double kmax(double a, double b) { return (a < b) ? b : a; }

/* IEEE-754 bit-level NaN / finiteness tests that survive -ffast-math (src/utils.h:9-72) */
static inline int ieee_isnan(double x)
{
    uint64_t b; memcpy(&b, &x, sizeof b);
    return (b & 0x7FF0000000000000ull) == 0x7FF0000000000000ull && (b & 0x000FFFFFFFFFFFFFull) != 0;
}
static inline int ieee_isfinite(double x)
{
    uint64_t b; memcpy(&b, &x, sizeof b);
    return (b & 0x7FF0000000000000ull) != 0x7FF0000000000000ull;
}

#if ORACLE_MD_ITER
/* indexing.cpp:113-130: the share of the last round of tiles that leaves threads idle */
static double o_overwork_factor(o_idx n, o_idx threads, o_idx tile)
{
    const double useful = (double) n / (double) tile;
    const double total = ceil(useful / (double) threads) * (double) threads;
    return 1 - useful / total;
}

/* indexing.cpp:137-191 (optimal_split_iter_space, tolerance 0.10): the largest tile of rows, from
 * `threads` rows down, whose overwork is zero, or else below the tolerance (a smaller one only if
 * it halves the overwork), or else the one of least overwork. */
static o_idx o_md_tile_rows(o_idx rows)
{
#ifdef _OPENMP
    const o_idx threads = omp_get_max_threads();
#else
    const o_idx threads = 1;
#endif
    const o_idx min_tile = threads < 4 ? threads : 4;
    double min_of = INFINITY, best_of = 0;
    o_idx optimal = 0, best = 0;
    for (o_idx tile = threads; tile > 0; tile--) {
        if (tile < min_tile) continue;
        const double of = o_overwork_factor(rows, threads, tile);
        if (of == 0) {
            best = tile;
            break;
        } else if (of < 0.10 && (best < tile || 2 * of < best_of)) {
            best = tile;
            best_of = of;
        } else if (of < min_of) {
            min_of = of;
            optimal = tile;
        }
    }
    if (best > 0) optimal = best;
    return optimal > 0 ? optimal : 1;
}
#endif

/* ------------------------------------------------------------------------- */
/* Indexing: indexing.h:39-73, indexing.cpp:5-110                            */
/* ------------------------------------------------------------------------- */

/* indexing.h:44 */
o_idx o_scale_index_1d(const OInnerRange1D* r, o_uidx i)
{
    return (o_idx) i * r->step + r->start;
}

/* indexing.h:67-72 */
o_idx o_scale_index_2d(const OInnerRange2D* r, o_uidx i)
{
    o_idx ix = (o_idx) (i / r->row_range_length);
    o_idx iy = (o_idx) (i % r->row_range_length);
    o_idx j = r->main_range_start + ix * r->main_range_step;
    return r->row_range_start + iy + j;
}

/* indexing.cpp:33-41, a Julia-style range length (step >= 0) */
static uint64_t range_length(int64_t start, int64_t step, int64_t stop)
{
    if ((start != stop) && ((step > 0) != (stop > start))) return 0;
    int64_t diff = stop - start;
    return (uint64_t) ((diff / step) + 1);
}

/* indexing.cpp:44-53 */
static int64_t range_last(int64_t start, int64_t step, int64_t stop)
{
    if (stop == start) return stop;
    int64_t remainder = (stop - start) % step;
    return stop - remainder;
}

uint64_t o_dr_col_length(const ODomainRange* d) { return range_length(d->col_start, d->col_step, d->col_end); } /* indexing.cpp:56-59 */
uint64_t o_dr_row_length(const ODomainRange* d) { return range_length(d->row_start, d->row_step, d->row_end); } /* indexing.cpp:62-65 */
uint64_t o_dr_length(const ODomainRange* d) { return o_dr_col_length(d) * o_dr_row_length(d); }                 /* indexing.cpp:68-71 */
int64_t  o_dr_begin(const ODomainRange* d) { return d->col_start + d->row_start; }                               /* indexing.cpp:101-104 */
int64_t  o_dr_end(const ODomainRange* d)                                                                         /* indexing.cpp:107-110 */
{
    return range_last(d->col_start, d->col_step, d->col_end) + range_last(d->row_start, d->row_step, d->row_end);
}

/* indexing.cpp:5-18 */
void o_dr_inflate_dir(ODomainRange* d, int axis, int64_t n)
{
    if (axis == O_AXIS_X) { d->row_start -= d->row_step * n; d->row_end += d->row_step * n; }
    else                  { d->col_start -= d->col_step * n; d->col_end += d->col_step * n; }
}

/* indexing.cpp:20-30 */
void o_dr_expand_dir(ODomainRange* d, int axis, int64_t n)
{
    if (axis == O_AXIS_X) d->row_end += d->row_step * n;
    else                  d->col_end += d->col_step * n;
}

/* indexing.cpp:83-89 */
void o_dr_iter1d(const ODomainRange* d, ORange* r, OInnerRange1D* ir)
{
    r->start = 0;
    r->end = (o_idx) (o_dr_end(d) - o_dr_begin(d) + 1);
    ir->start = (o_idx) o_dr_begin(d);
    ir->step = (o_idx) d->row_step;
}

/* indexing.cpp:74-80 */
void o_dr_direct_iter1d(const ODomainRange* d, ORange* r, OInnerRange1D* ir)
{
    r->start = 0;
    r->end = (o_idx) o_dr_length(d);
    ir->start = (o_idx) o_dr_begin(d);
    ir->step = (o_idx) d->row_step;
}

/* indexing.cpp:92-98 */
void o_dr_iter2d(const ODomainRange* d, ORange* r, OInnerRange2D* ir)
{
    r->start = 0;
    r->end = (o_idx) o_dr_length(d);
    ir->main_range_start = (o_idx) d->col_start;
    ir->main_range_step = (o_idx) d->col_step;
    ir->row_range_start = (o_idx) d->row_start;
    ir->row_range_length = o_dr_row_length(d);
}

/* ------------------------------------------------------------------------- */
/* Correctly rounded y^3 and y^4 (stand-in for Kokkos::pow(y, 3|4) =        */
/* glibc pow in the reference's IEEE build, fluid_eos.cpp:50-59).            */
/* Double-double products (exact via fma), rounded once at the end.          */
/* ------------------------------------------------------------------------- */

double armon_cr_pow3(double y)
{
    double p  = y * y;
    double e1 = fma(y, y, -p);      /* y^2 = p + e1 exactly */
    double q  = p * y;
    double e2 = fma(p, y, -q);      /* p*y = q + e2 exactly */
    double t  = e1 * y;             /* e1*y, relative error 2^-53 of a 2^-53 term */
    return q + (e2 + t);
}

double armon_cr_pow4(double y)
{
    double p  = y * y;
    double e1 = fma(y, y, -p);      /* y^2 = p + e1 exactly */
    double q  = p * p;
    double e2 = fma(p, p, -q);      /* p^2 = q + e2 exactly */
    double t  = (2.0 * p) * e1;     /* cross term 2*p*e1; e1^2 is below 2^-104 relative */
    return q + (e2 + t);
}

/* ------------------------------------------------------------------------- */
/* Kernels                                                                   */
/* ------------------------------------------------------------------------- */

/* acoustic.cpp:10-19 */
static inline void acoustic_Godunov(double rho_i, double rho_im, double c_i, double c_im,
                                    double u_i, double u_im, double p_i, double p_im,
                                    double* ustar_i, double* pstar_i)
{
    double rc_l = rho_im * c_im;
    double rc_r = rho_i * c_i;
    *ustar_i = (rc_l * u_im + rc_r * u_i + (p_im - p_i)) / (rc_l + rc_r);
    *pstar_i = (rc_r * p_im + rc_l * p_i + rc_l * rc_r * (u_im - u_i)) / (rc_l + rc_r);
}

/* acoustic.cpp:22-37 */
void o_acoustic(const ORange* range, const OInnerRange2D* ir, o_idx s,
                double* ustar, double* pstar,
                const double* rho, const double* u, const double* pmat, const double* cmat)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, i) {
        double us, ps;
        acoustic_Godunov(rho[i], rho[i - s], cmat[i], cmat[i - s], u[i], u[i - s], pmat[i], pmat[i - s], &us, &ps);
        ustar[i] = us;
        pstar[i] = ps;
    }
}

/* limiters.h:15-34 */
static inline double apply_limiter(int L, double r)
{
    switch (L) {
    case O_LIM_MINMOD:   return kmax(0.0, kmin(1.0, r));
    case O_LIM_SUPERBEE: return kmax(kmax(0.0, kmin(1.0, 2 * r)), kmin(2.0, r));
    default:             return 1.0;
    }
}

/* limiters.h:15-34, exported for the limiter-table tests */
double o_limiter(int L, double r) { return apply_limiter(L, r); }

/* acoustic.cpp:40-88 (template body) */
static void acoustic_GAD_impl(const ORange* range, const OInnerRange2D* ir, o_idx s, double dt, double dx,
                              double* ustar, double* pstar,
                              const double* rho, const double* u, const double* pmat, const double* cmat,
                              int L)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, i) {

        double ustar_im, pstar_im, ustar_i, pstar_i, ustar_ip, pstar_ip;
        acoustic_Godunov(rho[i - s], rho[i - 2 * s], cmat[i - s], cmat[i - 2 * s],
                         u[i - s], u[i - 2 * s], pmat[i - s], pmat[i - 2 * s], &ustar_im, &pstar_im);
        acoustic_Godunov(rho[i], rho[i - s], cmat[i], cmat[i - s],
                         u[i], u[i - s], pmat[i], pmat[i - s], &ustar_i, &pstar_i);
        acoustic_Godunov(rho[i + s], rho[i], cmat[i + s], cmat[i],
                         u[i + s], u[i], pmat[i + s], pmat[i], &ustar_ip, &pstar_ip);

        double r_um = (ustar_ip - u[i]) / (ustar_i - u[i - s] + 1e-6);
        double r_pm = (pstar_ip - pmat[i]) / (pstar_i - pmat[i - s] + 1e-6);
        double r_up = (u[i - s] - ustar_im) / (u[i] - ustar_i + 1e-6);
        double r_pp = (pmat[i - s] - pstar_im) / (pmat[i] - pstar_i + 1e-6);

        r_um = apply_limiter(L, r_um);
        r_pm = apply_limiter(L, r_pm);
        r_up = apply_limiter(L, r_up);
        r_pp = apply_limiter(L, r_pp);

        double dm_l = rho[i - s] * dx;
        double dm_r = rho[i] * dx;
        double Dm = (dm_l + dm_r) / 2;

        double rc_l = rho[i - s] * cmat[i - s];
        double rc_r = rho[i] * cmat[i];
        double theta = 0.5 * (1 - (rc_l + rc_r) / 2 * (dt / Dm));

        ustar[i] = ustar_i + theta * (r_up * (u[i] - ustar_i) - r_um * (ustar_i - u[i - s]));
        pstar[i] = pstar_i + theta * (r_pp * (pmat[i] - pstar_i) - r_pm * (pstar_i - pmat[i - s]));
    }
}

/* acoustic.cpp:91-108. Returns 0, or -1 for an invalid limiter (the
 * reference throws std::out_of_range through its exception handler). */
int o_acoustic_GAD(const ORange* range, const OInnerRange2D* ir, o_idx s, double dt, double dx,
                   double* ustar, double* pstar,
                   const double* rho, const double* u, const double* pmat, const double* cmat,
                   int limiter)
{
    if (limiter < O_LIM_NONE || limiter > O_LIM_SUPERBEE) return -1;
    acoustic_GAD_impl(range, ir, s, dt, dx, ustar, pstar, rho, u, pmat, cmat, limiter);
    return 0;
}

/* fluid_eos.cpp:9-23 */
void o_perfect_gas_EOS(const ORange* range, const OInnerRange2D* ir, double gamma,
                       const double* rho, const double* Emat, const double* umat, const double* vmat,
                       double* pmat, double* cmat, double* gmat)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, i) {
        double e = Emat[i] - 0.5 * (umat[i] * umat[i] + vmat[i] * vmat[i]);
        pmat[i] = (gamma - 1) * rho[i] * e;
        cmat[i] = sqrt(gamma * pmat[i] / rho[i]);
        gmat[i] = (1 + gamma) / 2;
    }
}

/* fluid_eos.cpp:26-62 */
void o_bizarrium_EOS(const ORange* range, const OInnerRange2D* ir,
                     const double* rho, const double* umat, const double* vmat, const double* Emat,
                     double* pmat, double* cmat, double* gmat)
{
    const double rho0 = 1e4, K0 = 1e11, Cv0 = 1e3, T0 = 300, eps0 = 0;
    const double G0 = 1.5, s = 1.5, q = -42080895. / 14941154., r = 727668333. / 149411540.;
    const OInnerRange2D ri = *ir;

    FOR_2D(range, ri, i) {

        double x = rho[i] / rho0 - 1;
        double g = G0 * (1 - rho0 / rho[i]);

        double f0 = (1 + (s / 3 - 2) * x + q * (x * x) + r * (x * x * x)) / (1 - s * x);
        double f1 = (s / 3 - 2 + 2 * q * x + 3 * r * (x * x) + s * f0) / (1 - s * x);
        double f2 = (2 * q + 6 * r * x + 2 * s * f1) / (1 - s * x);
        double f3 = (6 * r + 3 * s * f2) / (1 - s * x);

        double eps_k0 = eps0 - Cv0 * T0 * (1 + g) + 0.5 * (K0 / rho0) * (x * x) * f0;
        double pk0 = -Cv0 * T0 * G0 * rho0 + (0.5 * K0 * x * (1 + x) * (1 + x) * (2 * f0 + x * f1));
        double pk0prime = -0.5 * K0 * armon_cr_pow3(1 + x)
                * rho0 * (2 * (1 + 3 * x) * f0 + 2 * x * (2 + 3 * x) * f1 + (x * x) * (1 + x) * f2);
        double xp = x * (1 + x);
        double pk0second = 0.5 * K0 * armon_cr_pow4(1 + x) * (rho0 * rho0)
                * (12 * (1 + 2 * x) * f0 + 6 * (1 + 6 * x + 6 * (x * x)) * f1 + 6 * x * (1 + x) * (1 + 2 * x) * f2
                   + (xp * xp) * f3);

        double e = Emat[i] - 0.5 * (umat[i] * umat[i] + vmat[i] * vmat[i]);
        pmat[i] = pk0 + G0 * rho0 * (e - eps_k0);
        cmat[i] = sqrt(G0 * rho0 * (pmat[i] - pk0) - pk0prime) / rho[i];
        double G0rho0 = G0 * rho0;
        gmat[i] = 0.5 / (armon_cr_pow3(rho[i]) * (cmat[i] * cmat[i]))
                * (pk0second + (G0rho0 * G0rho0) * (pmat[i] - pk0));
    }
}

/* euler_projection.cpp:9-22 */
void o_cell_update(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                   const double* ustar, const double* pstar, double* rho, double* u, double* Emat)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, i) {
        double dm = rho[i] * dx;
        rho[i] = dm / (dx + dt * (ustar[i + s] - ustar[i]));
        u[i] += dt / dm * (pstar[i] - pstar[i + s]);
        Emat[i] += dt / dm * (pstar[i] * ustar[i] - pstar[i + s] * ustar[i + s]);
    }
}

/* euler_projection.cpp:25-48 */
void o_euler_projection(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                        const double* ustar, double* rho, double* umat, double* vmat, double* Emat,
                        const double* adv_rho, const double* adv_urho,
                        const double* adv_vrho, const double* adv_Erho)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, i) {

        double dX = dx + dt * (ustar[i + s] - ustar[i]);

        double tmp_rho  = (dX * rho[i]           - (adv_rho[i + s]  - adv_rho[i]))  / dx;
        double tmp_urho = (dX * rho[i] * umat[i] - (adv_urho[i + s] - adv_urho[i])) / dx;
        double tmp_vrho = (dX * rho[i] * vmat[i] - (adv_vrho[i + s] - adv_vrho[i])) / dx;
        double tmp_Erho = (dX * rho[i] * Emat[i] - (adv_Erho[i + s] - adv_Erho[i])) / dx;

        rho[i]  = tmp_rho;
        umat[i] = tmp_urho / tmp_rho;
        vmat[i] = tmp_vrho / tmp_rho;
        Emat[i] = tmp_Erho / tmp_rho;
    }
}

/* euler_projection.cpp:51-70 */
void o_advection_first_order(const ORange* range, const OInnerRange2D* ir, o_idx s, double dt,
                             const double* ustar, const double* rho, const double* umat,
                             const double* vmat, const double* Emat,
                             double* adv_rho, double* adv_urho, double* adv_vrho, double* adv_Erho)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, is) {
        o_idx i = is;   /* the donor shift below moves i, never the loop index */
        double disp = dt * ustar[i];
        i -= (o_idx) (disp > 0) * s;

        adv_rho[is]  = disp * (rho[i]);
        adv_urho[is] = disp * (rho[i] * umat[i]);
        adv_vrho[is] = disp * (rho[i] * vmat[i]);
        adv_Erho[is] = disp * (rho[i] * Emat[i]);
    }
}

/* euler_projection.cpp:73-79 */
static inline double slope_minmod(double u_im, double u_i, double u_ip, double r_m, double r_p)
{
    double D_u_p = r_p * (u_ip - u_i);
    double D_u_m = r_m * (u_i - u_im);
    double s = copysign(1.0, D_u_p);
    return s * kmax(0.0, kmin(s * D_u_p, s * D_u_m));
}

/* euler_projection.cpp:82-120 */
void o_advection_second_order(const ORange* range, const OInnerRange2D* ir, o_idx s, double dx, double dt,
                              const double* ustar, const double* rho, const double* umat,
                              const double* vmat, const double* Emat,
                              double* adv_rho, double* adv_urho, double* adv_vrho, double* adv_Erho)
{
    const OInnerRange2D r = *ir;
    FOR_2D(range, r, is) {
        o_idx i = is;   /* the donor shift below moves i, never the loop index */
        double disp = dt * ustar[i];
        double Dx;
        if (disp > 0) {
            Dx = -(dx - dt * ustar[i - s]);
            i = i - s;
        } else {
            Dx = dx + dt * ustar[i + s];
        }

        double Dx_lm = dx + dt * (ustar[i] - ustar[i - s]);
        double Dx_l  = dx + dt * (ustar[i + s] - ustar[i]);
        double Dx_lp = dx + dt * (ustar[i + 2 * s] - ustar[i + s]);

        double r_m = (2 * Dx_l) / (Dx_l + Dx_lm);
        double r_p = (2 * Dx_l) / (Dx_l + Dx_lp);

        double slope_r  = slope_minmod(rho[i - s], rho[i], rho[i + s], r_m, r_p);
        double slope_ur = slope_minmod(rho[i - s] * umat[i - s], rho[i] * umat[i], rho[i + s] * umat[i + s], r_m, r_p);
        double slope_vr = slope_minmod(rho[i - s] * vmat[i - s], rho[i] * vmat[i], rho[i + s] * vmat[i + s], r_m, r_p);
        double slope_Er = slope_minmod(rho[i - s] * Emat[i - s], rho[i] * Emat[i], rho[i + s] * Emat[i + s], r_m, r_p);

        double length_factor = Dx / (2 * Dx_l);
        adv_rho[is]  = disp * (rho[i] - slope_r * length_factor);
        adv_urho[is] = disp * (rho[i] * umat[i] - slope_ur * length_factor);
        adv_vrho[is] = disp * (rho[i] * vmat[i] - slope_vr * length_factor);
        adv_Erho[is] = disp * (rho[i] * Emat[i] - slope_Er * length_factor);
    }
}

/* halo_exchange.cpp:9-36 */
void o_boundary_conditions(const ORange* range, const OInnerRange1D* ir, o_idx stencil_width, o_idx disp,
                           double u_factor, double v_factor,
                           double* rho, double* umat, double* vmat, double* pmat,
                           double* cmat, double* gmat, double* Emat)
{
    const OInnerRange1D r = *ir;
    PARALLEL_FOR
    for (o_idx k = range->start; k < range->end; k++) {
        o_idx i = o_scale_index_1d(&r, (o_uidx) k);
        o_idx ip = i + disp;
        for (o_idx w = 0; w < stencil_width; w++) {
            rho[i]  = rho[ip];
            umat[i] = umat[ip] * u_factor;
            vmat[i] = vmat[ip] * v_factor;
            pmat[i] = pmat[ip];
            cmat[i] = cmat[ip];
            gmat[i] = gmat[ip];
            Emat[i] = Emat[ip];
            i  -= disp;
            ip += disp;
        }
    }
}

/* halo_exchange.cpp:39-64 */
void o_read_border_array(const ORange* range, const OInnerRange2D* ir, o_idx side_length, o_idx nghost,
                         const double* rho, const double* umat, const double* vmat, const double* pmat,
                         const double* cmat, const double* gmat, const double* Emat,
                         double* value_array)
{
    const OInnerRange2D r = *ir;
    PARALLEL_FOR
    for (o_idx k = range->start; k < range->end; k++) {
        o_idx idx = o_scale_index_2d(&r, (o_uidx) k);
        o_idx itr = k;
        o_idx i = itr / nghost;
        o_idx i_g = itr % nghost;
        o_idx i_arr = (i_g * side_length + i) * 7;
        value_array[i_arr + 0] = rho[idx];
        value_array[i_arr + 1] = umat[idx];
        value_array[i_arr + 2] = vmat[idx];
        value_array[i_arr + 3] = pmat[idx];
        value_array[i_arr + 4] = cmat[idx];
        value_array[i_arr + 5] = gmat[idx];
        value_array[i_arr + 6] = Emat[idx];
    }
}

/* halo_exchange.cpp:67-92 */
void o_write_border_array(const ORange* range, const OInnerRange2D* ir, o_idx side_length, o_idx nghost,
                          double* rho, double* umat, double* vmat, double* pmat,
                          double* cmat, double* gmat, double* Emat,
                          const double* value_array)
{
    const OInnerRange2D r = *ir;
    PARALLEL_FOR
    for (o_idx k = range->start; k < range->end; k++) {
        o_idx idx = o_scale_index_2d(&r, (o_uidx) k);
        o_idx itr = k;
        o_idx i = itr / nghost;
        o_idx i_g = itr % nghost;
        o_idx i_arr = (i_g * side_length + i) * 7;
        rho[idx]  = value_array[i_arr + 0];
        umat[idx] = value_array[i_arr + 1];
        vmat[idx] = value_array[i_arr + 2];
        pmat[idx] = value_array[i_arr + 3];
        cmat[idx] = value_array[i_arr + 4];
        gmat[idx] = value_array[i_arr + 5];
        Emat[idx] = value_array[i_arr + 6];
    }
}

/* init_test.cpp:12-48 and test_cases.h:16-121: per-test initial states */
void o_reference_test_state(int test, double r, OTestState* t)
{
    memset(t, 0, sizeof(*t));
    switch (test) {
    case O_TEST_BIZARRIUM:                                   /* init_test.cpp:25-35 */
        t->high_rho = 1.42857142857e+4; t->low_rho = 10000.;
        t->high_E = 4.48657821135e+6;   t->low_E = 31250;
        t->high_u = 0.;                 t->low_u = 250.;
        break;
    case O_TEST_SEDOV:                                       /* init_test.cpp:38-48 */
        t->high_rho = 1.; t->low_rho = 1.;
        t->high_E = 0.851072 / (M_PI * (r * r));
        t->low_E = 2.5e-14;
        break;
    default:                                                 /* init_test.cpp:12-22 (Sod, Sod_y, Sod_circ) */
        t->high_rho = 1.;  t->low_rho = 0.125;
        t->high_E = 2.5;   t->low_E = 2.0;
        break;
    }
}

/* test_cases.h:50-51, 61-62, 72-76, 94-95, 112-116 */
static inline int test_region_high(int test, double r, double x, double y)
{
    switch (test) {
    case O_TEST_SOD:       return x <= 0.5;
    case O_TEST_SOD_Y:     return y <= 0.5;
    case O_TEST_SOD_CIRC:  return ((x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5)) <= 0.09;
    case O_TEST_BIZARRIUM: return x <= 0.5;
    case O_TEST_SEDOV:     return (x * x + y * y) <= r * r;
    default:               return 0;
    }
}

/* init_test.cpp:51-104 (template body) and :107-143 (dispatch). Returns -1
 * for an invalid test index (the reference throws std::out_of_range). */
/* init_test.cpp:51-94 with the high/low states given by the caller (o_init_test passes the
 * reference's states of `test`; the batched ensemble passes perturbed ones). */
int o_init_test_states(const ORange* range, const OInnerRange2D* ir,
                       o_idx row_length, o_idx nb_ghosts, o_idx nx, o_idx ny, double sx, double sy,
                       double ox, double oy, o_idx pos_x, o_idx pos_y, o_idx g_nx, o_idx g_ny,
                       double* x, double* y, double* rho, double* Emat, double* umat, double* vmat,
                       double* domain_mask, double* pmat, double* cmat, double* gmat,
                       double* ustar, double* pstar, int test, bool debug_indexes, double test_option,
                       const OTestState* states)
{
    if (test < O_TEST_SOD || test > O_TEST_SEDOV) return -1;
    const OTestState tp = *states;
    const OInnerRange2D r = *ir;

    FOR_2D(range, r, i) {

        o_idx ix = (i % row_length) - nb_ghosts;
        o_idx iy = (i / row_length) - nb_ghosts;
        o_idx g_ix = ix + pos_x;
        o_idx g_iy = iy + pos_y;

        x[i] = (double) g_ix / (double) nx * sx + ox;
        y[i] = (double) g_iy / (double) ny * sy + oy;

        double x_mid = x[i] + sx / (double) (2 * g_nx);
        double y_mid = y[i] + sy / (double) (2 * g_ny);

        if (debug_indexes) {
            rho[i] = (double) (i + 1);
            Emat[i] = (double) (i + 1);
            umat[i] = (double) (i + 1);
            vmat[i] = (double) (i + 1);
        } else if (test_region_high(test, test_option, x_mid, y_mid)) {
            rho[i] = tp.high_rho; Emat[i] = tp.high_E; umat[i] = tp.high_u; vmat[i] = tp.high_v;
        } else {
            rho[i] = tp.low_rho;  Emat[i] = tp.low_E;  umat[i] = tp.low_u;  vmat[i] = tp.low_v;
        }

        domain_mask[i] = (0 <= ix && ix < nx && 0 <= iy && iy < ny);

        pmat[i] = 0;
        cmat[i] = 1;
        gmat[i] = 0;
        ustar[i] = 0;
        pstar[i] = 0;
    }
    return 0;
}

int o_init_test(const ORange* range, const OInnerRange2D* ir,
                o_idx row_length, o_idx nb_ghosts, o_idx nx, o_idx ny, double sx, double sy,
                double ox, double oy, o_idx pos_x, o_idx pos_y, o_idx g_nx, o_idx g_ny,
                double* x, double* y, double* rho, double* Emat, double* umat, double* vmat,
                double* domain_mask, double* pmat, double* cmat, double* gmat,
                double* ustar, double* pstar, int test, bool debug_indexes, double test_option)
{
    if (test < O_TEST_SOD || test > O_TEST_SEDOV) return -1;
    OTestState tp;
    o_reference_test_state(test, test_option, &tp);
    return o_init_test_states(range, ir, row_length, nb_ghosts, nx, ny, sx, sy, ox, oy, pos_x, pos_y, g_nx, g_ny,
                              x, y, rho, Emat, umat, vmat, domain_mask, pmat, cmat, gmat, ustar, pstar, test,
                              debug_indexes, test_option, &tp);
}

/* reductions.cpp:8-24. Min is order independent, so threads do not change the result.
 * The reduction starts from Kokkos' identity for Min on double, the largest finite value
 * (the INFINITY initialiser of reductions.cpp:12 is overwritten by the reducer's init):
 * a range where every cell bound is +inf yields DBL_MAX. */
double o_dt_CFL(const ORange* range, const OInnerRange1D* ir, double dx, double dy,
                const double* umat, const double* vmat, const double* cmat, const double* domain_mask)
{
    const OInnerRange1D r = *ir;
    double dt = DBL_MAX;
#pragma omp parallel
    {
        double dt_loop = DBL_MAX;
#pragma omp for schedule(static) nowait
        for (o_idx k = range->start; k < range->end; k++) {
            o_idx i = o_scale_index_1d(&r, (o_uidx) k);
            double max_cx = kmax(fabs(umat[i] + cmat[i]), fabs(umat[i] - cmat[i])) * domain_mask[i];
            double max_cy = kmax(fabs(vmat[i] + cmat[i]), fabs(vmat[i] - cmat[i])) * domain_mask[i];
            dt_loop = kmin(dt_loop, kmin(dx / max_cx, dy / max_cy));
        }
#pragma omp critical
        {
            /* Kokkos::Min join: dest = (src < dest) ? src : dest */
            dt = (dt_loop < dt) ? dt_loop : dt;
        }
    }
    return dt;
}

/* Neumaier compensated accumulation */
typedef struct { double s, c; } OSum;
static inline void osum_add(OSum* a, double v)
{
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
    else                       a->c += (v - t) + a->s;
    a->s = t;
}

/* reductions.cpp:27-49. Fixed chunking (independent of the thread count),
 * compensated inside chunks and across chunks. */
void o_conservation_vars(const ORange* range, const OInnerRange1D* ir, double dx,
                         const double* rho, const double* Emat, const double* domain_mask,
                         double* total_mass, double* total_energy)
{
    const OInnerRange1D r = *ir;
    const double ds = dx * dx;
    const o_idx n = range->end - range->start;
    const o_idx chunk = 4096;
    const o_idx nchunks = (n + chunk - 1) / chunk;
    OSum* parts = (OSum*) calloc((size_t) (2 * (nchunks > 0 ? nchunks : 1)), sizeof(OSum));

    PARALLEL_FOR
    for (o_idx c = 0; c < nchunks; c++) {
        OSum m = {0, 0}, e = {0, 0};
        o_idx k0 = range->start + c * chunk;
        o_idx k1 = k0 + chunk < range->end ? k0 + chunk : range->end;
        for (o_idx k = k0; k < k1; k++) {
            o_idx i = o_scale_index_1d(&r, (o_uidx) k);
            double cell_mass = rho[i] * domain_mask[i] * ds;
            double cell_energy = cell_mass * Emat[i];
            osum_add(&m, cell_mass);
            osum_add(&e, cell_energy);
        }
        parts[2 * c] = m;
        parts[2 * c + 1] = e;
    }

    OSum m = {0, 0}, e = {0, 0};
    for (o_idx c = 0; c < nchunks; c++) {
        osum_add(&m, parts[2 * c].s); osum_add(&m, parts[2 * c].c);
        osum_add(&e, parts[2 * c + 1].s); osum_add(&e, parts[2 * c + 1].c);
    }
    free(parts);
    *total_mass = m.s + m.c;
    *total_energy = e.s + e.c;
}

/* ------------------------------------------------------------------------- */
/* Parameters: parameters.h:34-108, parameters.cpp                           */
/* ------------------------------------------------------------------------- */

void o_params_defaults(OParams* p)
{
    memset(p, 0, sizeof(*p));
    p->test = O_TEST_SOD;
    p->scheme = O_SCHEME_GAD;
    p->riemann = 0;
    p->projection = O_PROJ_EULER;
    p->limiter = O_LIM_MINMOD;
    p->current_axis = O_AXIS_X;
    p->axis_splitting = O_SPLIT_SEQUENTIAL;
    p->nb_ghosts = 3;
    p->nx = 10;
    p->ny = 10;
    p->stencil_width = -1;
    p->domain_size[0] = 0;
    p->domain_size[1] = 0;
    p->domain_origin[0] = NAN;
    p->domain_origin[1] = NAN;
    p->max_cycles = 100;
    p->max_time = 0;
    p->verbose = 2;
}

/* parameters.cpp:11-18 */
int o_min_ghosts(const OParams* p)
{
    int m = 1;
    m += p->scheme == O_SCHEME_GAD;
    m += p->projection == O_PROJ_EULER_2ND;
    m += (p->projection == O_PROJ_EULER_2ND) && (p->scheme == O_SCHEME_GAD);
    return m;
}

/* test_cases.h:43-44, 82-83, 108-109 */
double o_test_default_cfl(int test)
{
    switch (test) {
    case O_TEST_BIZARRIUM: return 0.6;
    case O_TEST_SEDOV:     return 0.7;
    default:               return 0.95;
    }
}

double o_test_default_max_time(int test)
{
    switch (test) {
    case O_TEST_BIZARRIUM: return 80e-6;
    case O_TEST_SEDOV:     return 1.0;
    default:               return 0.20;
    }
}

/* test_cases.h:36-38 (BC_FreeFlow, BC_Dirichlet_X/Y) and the per-test
 * boundaryCondition overrides (:45-48, :56-59, :67-70, :84-92, :110) */
void o_test_bc(int test, int side, double* uf, double* vf)
{
    const int lr = (side == O_SIDE_LEFT || side == O_SIDE_RIGHT);
    double u = 1, v = 1;
    switch (test) {
    case O_TEST_SOD:       if (lr) u = -1; break;
    case O_TEST_SOD_Y:     if (!lr) v = -1; break;
    case O_TEST_SOD_CIRC:  if (lr) u = -1; else v = -1; break;
    case O_TEST_BIZARRIUM:
        if (side == O_SIDE_LEFT) u = -1;
        else if (side == O_SIDE_RIGHT) { /* free flow */ }
        else v = -1;
        break;
    default: break;         /* Sedov: free flow everywhere */
    }
    *uf = u;
    *vf = v;
}

/* parameters.cpp:52-67 */
static void init_indexing(OParams* p)
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

/* parameters.cpp:21-49 */
void o_params_init(OParams* p)
{
    if (p->cfl == 0) p->cfl = o_test_default_cfl(p->test);
    if (p->max_time == 0) p->max_time = o_test_default_max_time(p->test);
    if (p->domain_size[0] <= 0 || p->domain_size[1] <= 0) {
        if (p->test == O_TEST_SEDOV) { p->domain_size[0] = 2; p->domain_size[1] = 2; }
        else                         { p->domain_size[0] = 1; p->domain_size[1] = 1; }
    }
    if (ieee_isnan(p->domain_origin[0]) || ieee_isnan(p->domain_origin[1])) {
        if (p->test == O_TEST_SEDOV) { p->domain_origin[0] = -1; p->domain_origin[1] = -1; }
        else                         { p->domain_origin[0] = 0;  p->domain_origin[1] = 0; }
    }
    if (p->stencil_width == -1) p->stencil_width = o_min_ghosts(p);
    init_indexing(p);
    o_update_axis(p, p->current_axis);
}

/* parameters.cpp:70-109 (messages go to stderr like the reference) */
int o_params_check(const OParams* p)
{
    if (p->projection == O_PROJ_NONE) { fprintf(stderr, "ERROR: Lagrangian mode is not supported\n"); return 0; }
    if (p->nx <= 0 || p->ny <= 0) { fprintf(stderr, "ERROR: One of the dimensions of the domain is 0 or negative\n"); return 0; }
    if (p->cst_dt && p->Dt == 0.) { fprintf(stderr, "ERROR: Constant time step is set but the initial time step is 0\n"); return 0; }
    int mg = o_min_ghosts(p);
    if (p->nb_ghosts < mg) fprintf(stderr, "WARNING: There is not enough ghost cells for the given parameters.\n");
    if (p->stencil_width < mg) fprintf(stderr, "WARNING: The stencil width given does not cover the estimated width.\n");
    if (p->stencil_width > p->nb_ghosts) { fprintf(stderr, "ERROR: The stencil width cannot be bigger than the number of ghost cells\n"); return 0; }
    return 1;
}

/* parameters.cpp:195-210 */
void o_update_axis(OParams* p, int axis)
{
    p->current_axis = axis;
    if (axis == O_AXIS_X) { p->s = 1;             p->dx = p->domain_size[0] / (double) p->nx; }
    else                  { p->s = p->row_length; p->dx = p->domain_size[1] / (double) p->ny; }
}

/* parameters.cpp:212-255 */
int o_split_axes(const OParams* p, int axes[3], double factors[3])
{
    int a1, a2;
    if (p->cycle % 2 == 0) { a1 = O_AXIS_Y; a2 = O_AXIS_X; }
    else                   { a1 = O_AXIS_X; a2 = O_AXIS_Y; }
    switch (p->axis_splitting) {
    case O_SPLIT_SEQUENTIAL:
        axes[0] = O_AXIS_X; factors[0] = 1.0; axes[1] = O_AXIS_Y; factors[1] = 1.0; return 2;
    case O_SPLIT_SEQUENTIAL_SYM:
        axes[0] = a1; factors[0] = 1.0; axes[1] = a2; factors[1] = 1.0; return 2;
    case O_SPLIT_STRANG:
        axes[0] = a1; factors[0] = 0.5; axes[1] = a2; factors[1] = 1.0; axes[2] = a1; factors[2] = 0.5; return 3;
    case O_SPLIT_X_ONLY:
        axes[0] = O_AXIS_X; factors[0] = 1.0; return 1;
    case O_SPLIT_Y_ONLY:
        axes[0] = O_AXIS_Y; factors[0] = 1.0; return 1;
    default:
        return 0;
    }
}

/* ------------------------------------------------------------------------- */
/* Iteration domains: ranges.cpp                                             */
/* ------------------------------------------------------------------------- */

int64_t o_index_1D(const OParams* p, int i, int j) { return p->index_start + (int64_t) j * p->idx_row + (int64_t) i * p->idx_col; }

ODomainRange o_real_domain(const OParams* p)
{
    ODomainRange d = { o_index_1D(p, 0, 0), p->row_length, o_index_1D(p, 0, p->ny - 1), 0, 1, p->nx - 1 };
    return d;
}

ODomainRange o_domain_fluxes(const OParams* p)
{
    ODomainRange d = o_real_domain(p);
    int64_t extra = 1 + (p->projection == O_PROJ_EULER_2ND);
    o_dr_inflate_dir(&d, p->current_axis, extra);
    o_dr_expand_dir(&d, p->current_axis, 1);
    return d;
}

ODomainRange o_domain_cell_update(const OParams* p)
{
    ODomainRange d = o_real_domain(p);
    int64_t extra = 1 + (p->projection == O_PROJ_EULER_2ND);
    o_dr_inflate_dir(&d, p->current_axis, extra);
    return d;
}

ODomainRange o_domain_advection(const OParams* p)
{
    ODomainRange d = o_real_domain(p);
    o_dr_expand_dir(&d, p->current_axis, 1);
    return d;
}

ODomainRange o_complete_domain(const OParams* p)
{
    ODomainRange d = o_real_domain(p);
    o_dr_inflate_dir(&d, O_AXIS_X, p->nb_ghosts);
    o_dr_inflate_dir(&d, O_AXIS_Y, p->nb_ghosts);
    return d;
}

ODomainRange o_boundary_conditions_domain(const OParams* p, int side, int* disp)
{
    int i_start = 0, stride = 1, loop_range = 0;
    switch (side) {
    case O_SIDE_LEFT:   stride = p->row_length; i_start = (int) o_index_1D(p, -1, 0);    loop_range = p->ny; *disp = 1;              break;
    case O_SIDE_RIGHT:  stride = p->row_length; i_start = (int) o_index_1D(p, p->nx, 0); loop_range = p->ny; *disp = -1;             break;
    case O_SIDE_TOP:    stride = 1;             i_start = (int) o_index_1D(p, 0, p->ny); loop_range = p->nx; *disp = -p->row_length; break;
    default:            stride = 1;             i_start = (int) o_index_1D(p, 0, -1);    loop_range = p->nx; *disp = p->row_length;  break;
    }
    ODomainRange d = { 0, 1, 0, i_start, stride, (int64_t) i_start + (int64_t) loop_range * stride - stride };
    return d;
}

/* ------------------------------------------------------------------------- */
/* Data (data.h:46-64)                                                       */
/* ------------------------------------------------------------------------- */

int o_data_alloc(OData* d, int64_t n)
{
    memset(d, 0, sizeof(*d));
    double** arrs[] = { &d->x, &d->y, &d->rho, &d->umat, &d->vmat, &d->Emat, &d->pmat, &d->cmat, &d->gmat,
                        &d->ustar, &d->pstar, &d->work_array_1, &d->work_array_2, &d->work_array_3,
                        &d->work_array_4, &d->domain_mask };
    size_t bytes = (size_t) (n > 0 ? n : 1) * sizeof(double);
    bytes = (bytes + 63) / 64 * 64;
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) {
        *arrs[k] = (double*) aligned_alloc(64, bytes);
        if (!*arrs[k]) { o_data_free(d); return -1; }
        /* the reference allocates without initialisation; fill with NaN so
         * that any read of a never-written cell shows up */
        double* a = *arrs[k];
        PARALLEL_FOR
        for (int64_t i = 0; i < n; i++) a[i] = NAN;
    }
    return 0;
}

void o_data_free(OData* d)
{
    double* arrs[] = { d->x, d->y, d->rho, d->umat, d->vmat, d->Emat, d->pmat, d->cmat, d->gmat,
                       d->ustar, d->pstar, d->work_array_1, d->work_array_2, d->work_array_3,
                       d->work_array_4, d->domain_mask };
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); k++) free(arrs[k]);
    memset(d, 0, sizeof(*d));
}

/* ------------------------------------------------------------------------- */
/* Driver: armon_2D.cpp                                                      */
/* ------------------------------------------------------------------------- */

/* armon_2D.cpp:70-82 */
void o_numerical_fluxes(const OParams* p, OData* d)
{
    const double* u = p->current_axis == O_AXIS_X ? d->umat : d->vmat;
    ODomainRange dr = o_domain_fluxes(p);
    ORange r; OInnerRange2D ir;
    o_dr_iter2d(&dr, &r, &ir);
    if (p->scheme == O_SCHEME_GODUNOV)
        o_acoustic(&r, &ir, p->s, d->ustar, d->pstar, d->rho, u, d->pmat, d->cmat);
    else
        o_acoustic_GAD(&r, &ir, p->s, p->cycle_dt, p->dx, d->ustar, d->pstar, d->rho, u, d->pmat, d->cmat, p->limiter);
}

/* armon_2D.cpp:85-101 */
void o_update_EOS(const OParams* p, OData* d)
{
    ODomainRange dr = o_real_domain(p);
    ORange r; OInnerRange2D ir;
    o_dr_iter2d(&dr, &r, &ir);
    if (p->test == O_TEST_BIZARRIUM)
        o_bizarrium_EOS(&r, &ir, d->rho, d->umat, d->vmat, d->Emat, d->pmat, d->cmat, d->gmat);
    else
        o_perfect_gas_EOS(&r, &ir, 1.4, d->rho, d->Emat, d->umat, d->vmat, d->pmat, d->cmat, d->gmat);
}

/* armon_2D.cpp:104-109 */
void o_cell_update_params(const OParams* p, OData* d)
{
    ODomainRange dr = o_domain_cell_update(p);
    ORange r; OInnerRange2D ir;
    o_dr_iter2d(&dr, &r, &ir);
    double* u = p->current_axis == O_AXIS_X ? d->umat : d->vmat;
    o_cell_update(&r, &ir, p->s, p->dx, p->cycle_dt, d->ustar, d->pstar, d->rho, u, d->Emat);
}

/* test_cases.h:117-119: only Sedov has a test option (armon_2D.cpp:116-118) */
double o_test_option(const OParams* p)
{
    double dx = p->domain_size[0] / (double) p->nx;
    double dy = p->domain_size[1] / (double) p->ny;
    return p->test == O_TEST_SEDOV ? hypot(dx, dy) / M_SQRT2 : 0.;
}

/* armon_2D.cpp:112-127 */
void o_init_test_params(const OParams* p, OData* d, bool debug_indexes)
{
    ODomainRange dr = o_complete_domain(p);
    ORange r; OInnerRange2D ir;
    o_dr_iter2d(&dr, &r, &ir);
    o_init_test(&r, &ir, p->row_length, p->nb_ghosts, p->nx, p->ny,
                p->domain_size[0], p->domain_size[1], p->domain_origin[0], p->domain_origin[1],
                0, 0, p->nx, p->ny,
                d->x, d->y, d->rho, d->Emat, d->umat, d->vmat, d->domain_mask,
                d->pmat, d->cmat, d->gmat, d->ustar, d->pstar, p->test, debug_indexes, o_test_option(p));
}

/* o_init_test_params with caller-given high/low states (a batched-ensemble domain). */
void o_init_test_params_states(const OParams* p, OData* d, const OTestState* states)
{
    ODomainRange dr = o_complete_domain(p);
    ORange r; OInnerRange2D ir;
    o_dr_iter2d(&dr, &r, &ir);
    o_init_test_states(&r, &ir, p->row_length, p->nb_ghosts, p->nx, p->ny,
                       p->domain_size[0], p->domain_size[1], p->domain_origin[0], p->domain_origin[1],
                       0, 0, p->nx, p->ny,
                       d->x, d->y, d->rho, d->Emat, d->umat, d->vmat, d->domain_mask,
                       d->pmat, d->cmat, d->gmat, d->ustar, d->pstar, p->test, false, o_test_option(p), states);
}

/* armon_2D.cpp:130-140 */
static void boundary_conditions_side(const OParams* p, OData* d, int side)
{
    int disp = 0;
    ODomainRange dr = o_boundary_conditions_domain(p, side, &disp);
    double uf, vf;
    if (p->bc_override) {
        uf = p->bc_factors[side][0];
        vf = p->bc_factors[side][1];
    } else {
        o_test_bc(p->test, side, &uf, &vf);
    }
    ORange r; OInnerRange1D ir;
    o_dr_direct_iter1d(&dr, &r, &ir);
    o_boundary_conditions(&r, &ir, p->stencil_width, disp, uf, vf,
                          d->rho, d->umat, d->vmat, d->pmat, d->cmat, d->gmat, d->Emat);
}

/* armon_2D.cpp:143-151 */
void o_boundary_conditions_all(const OParams* p, OData* d)
{
    if (p->current_axis == O_AXIS_X) {
        boundary_conditions_side(p, d, O_SIDE_LEFT);
        boundary_conditions_side(p, d, O_SIDE_RIGHT);
    } else {
        boundary_conditions_side(p, d, O_SIDE_TOP);
        boundary_conditions_side(p, d, O_SIDE_BOTTOM);
    }
}

/* armon_2D.cpp:188-209 */
void o_projection_remap(const OParams* p, OData* d)
{
    if (p->projection == O_PROJ_NONE) return;
    ORange r; OInnerRange2D ir;
    ODomainRange da = o_domain_advection(p);
    o_dr_iter2d(&da, &r, &ir);
    if (p->projection == O_PROJ_EULER)
        o_advection_first_order(&r, &ir, p->s, p->cycle_dt, d->ustar, d->rho, d->umat, d->vmat, d->Emat,
                                d->work_array_1, d->work_array_2, d->work_array_3, d->work_array_4);
    else
        o_advection_second_order(&r, &ir, p->s, p->dx, p->cycle_dt, d->ustar, d->rho, d->umat, d->vmat, d->Emat,
                                 d->work_array_1, d->work_array_2, d->work_array_3, d->work_array_4);
    ODomainRange dr = o_real_domain(p);
    o_dr_iter2d(&dr, &r, &ir);
    o_euler_projection(&r, &ir, p->s, p->dx, p->cycle_dt, d->ustar, d->rho, d->umat, d->vmat, d->Emat,
                       d->work_array_1, d->work_array_2, d->work_array_3, d->work_array_4);
}

/* armon_2D.cpp:212-226 */
double o_local_time_step(const OParams* p, OData* d, double prev_dt)
{
    double dx = p->domain_size[0] / (double) p->nx;
    double dy = p->domain_size[1] / (double) p->ny;
    ODomainRange dr = o_real_domain(p);
    ORange r; OInnerRange1D ir;
    o_dr_iter1d(&dr, &r, &ir);
    double dt = o_dt_CFL(&r, &ir, dx, dy, d->umat, d->vmat, d->cmat, d->domain_mask);
    if (!ieee_isfinite(dt) || dt <= 0) return dt;
    else if (prev_dt == 0) return p->cfl * dt;
    else return kmin(p->cfl * dt, 1.05 * prev_dt);
}

/* armon_2D.cpp:229-252. Returns 1 where the reference exits on an invalid dt. */
int o_time_step(OParams* p, OData* d)
{
    p->current_cycle_dt = p->next_cycle_dt;
    if (p->cst_dt) {
        p->next_cycle_dt = p->Dt;
    } else if (!p->dt_on_even_cycles || (p->cycle % 2 == 0) || p->current_cycle_dt == 0) {
        double local_dt = o_local_time_step(p, d, p->current_cycle_dt);
        if (!ieee_isfinite(local_dt) || local_dt <= 0) return 1;
        p->next_cycle_dt = local_dt;
    } else {
        p->next_cycle_dt = p->current_cycle_dt;
    }
    if (p->cycle == 0) p->current_cycle_dt = p->next_cycle_dt;
    return 0;
}

/* armon_2D.cpp:304-324 */
int o_solver_cycle(OParams* p, OData* d)
{
    if (o_time_step(p, d)) return 1;
    int axes[3]; double factors[3];
    int n = o_split_axes(p, axes, factors);
    for (int k = 0; k < n; k++) {
        o_update_axis(p, axes[k]);
        p->cycle_dt = p->current_cycle_dt * factors[k];
        o_update_EOS(p, d);
        o_boundary_conditions_all(p, d);
        o_numerical_fluxes(p, d);
        o_cell_update_params(p, d);
        o_projection_remap(p, d);
    }
    return 0;
}

/* armon_2D.cpp:287-298 */
void o_conservation_vars_params(const OParams* p, OData* d, double* mass, double* energy)
{
    ODomainRange dr = o_real_domain(p);
    ORange r; OInnerRange1D ir;
    o_dr_iter1d(&dr, &r, &ir);
    o_conservation_vars(&r, &ir, p->dx, d->rho, d->Emat, d->domain_mask, mass, energy);
}

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

/* armon_2D.cpp:327-386 */
int o_time_loop(OParams* p, OData* d, double* loop_time, double* final_dt, int* cycles)
{
    int status = 0;
    p->cycle = 0;
    p->time = 0.;
    p->current_cycle_dt = 0.;
    p->next_cycle_dt = 0.;

    double t0 = now_seconds();
    o_update_axis(p, O_AXIS_X);
    o_update_EOS(p, d);

    double initial_mass = 0, initial_energy = 0;
    if (p->verbose <= 1) o_conservation_vars_params(p, d, &initial_mass, &initial_energy);

    while (p->time < p->max_time && p->cycle < p->max_cycles) {
        if (o_solver_cycle(p, d)) { status = 1; break; }
        if (p->verbose <= 1) {
            double m, e;
            o_conservation_vars_params(p, d, &m, &e);
            printf("Cycle = %4d, dt = %.18f, t = %.18f, |dM| = %#8.6g%%, |dE| = %#8.6g%%\n",
                   p->cycle + 1, p->current_cycle_dt, p->time,
                   fabs(initial_mass - m) / initial_mass * 100, fabs(initial_energy - e) / initial_energy * 100);
        }
        p->time += p->current_cycle_dt;
        p->cycle++;
    }
    double t1 = now_seconds();
    if (loop_time) *loop_time = t1 - t0;
    if (final_dt) *final_dt = p->next_cycle_dt;
    if (cycles) *cycles = p->cycle;
    return status;
}

int o_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void o_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void) n;
#endif
}

int o_is_strict(void)
{
#ifdef __FAST_MATH__
    return 0;
#else
    return 1;
#endif
}
