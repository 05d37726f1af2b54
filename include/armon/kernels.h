/*
 * Armon kernel C-ABI — MI355X (gfx950) implementation.
 *
 * Drop-in replacement for the reference's src/kernels/kernels.h:10-86: the same
 * 14 entry points, argument order and meaning, plus the ABI helpers of
 * src/kernels/utils.cpp:16-27. Behind them run hand-written HIP kernels.
 *
 * Differences a caller must know:
 *   - `view` is a POD handle {device pointer, extent, label} instead of
 *     Kokkos::View<flt_t*> (common.h:24), whose layout is Kokkos-internal.
 *     The caller owns the memory; kernels never allocate, free or retain it.
 *   - Idx/UIdx are fixed to int64_t/uint64_t (reported by idx_size() and
 *     is_uidx_signed()).
 *   - Every entry point checks, on the host, that each index it will touch
 *     lies in [0, view.extent) and, when a label is set, that the view order
 *     matches (utils.h:74-96). Failures go through raise_exception_handler and
 *     the entry point returns without launching.
 *   - Launches are asynchronous on the library stream (armon_set_stream);
 *     dt_CFL and conservation_vars block and return host values, as in the
 *     reference (reductions.cpp:8-49).
 *
 * In C++ the range and output arguments are references, exactly as in the
 * reference header; in C they are pointers. Both have the same ABI.
 */
#ifndef ARMON_HIP_KERNELS_H
#define ARMON_HIP_KERNELS_H

#include <stdbool.h>
#include <stdint.h>

#if defined(__GNUC__)
#define ARMON_EXPORT __attribute__((visibility("default")))
#else
#define ARMON_EXPORT
#endif

#ifdef __cplusplus
#define ARMON_CREF(T) const T&
#define ARMON_REF(T) T&
extern "C" {
#else
#define ARMON_CREF(T) const T*
#define ARMON_REF(T) T*
#endif

/* common.h:8-12: flt_t is double, or float in the reference's single-precision build
 * (USE_SINGLE_PRECISION). The two builds are two libraries: lib/libarmon.so (FP64, with the
 * solver extensions of armon.h) and lib/libarmon_f32.so (flt_t = float: the kernel entry points
 * of this header and the memory plumbing of armon.h, no solver). A caller that defines
 * ARMON_SINGLE_PRECISION links libarmon_f32.so; flt_size() reports which one is loaded. */
#if defined(ARMON_SINGLE_PRECISION) && ARMON_SINGLE_PRECISION
typedef float flt_t;
#else
typedef double flt_t;
#endif

/* indexing.h:14-19 */
typedef int64_t Idx;
typedef uint64_t UIdx;

/* indexing.h:28-33: open interval [start, end) of linear indices */
typedef struct Range { Idx start; Idx end; } Range;

/* indexing.h:39-45: array index = i * step + start */
typedef struct InnerRange1D { Idx start; Idx step; } InnerRange1D;

/* indexing.h:60-73: array index = row_range_start + (i % row_range_length)
 *                                + main_range_start + (i / row_range_length) * main_range_step */
typedef struct InnerRange2D {
    Idx main_range_start;
    Idx main_range_step;
    Idx row_range_start;
    UIdx row_range_length;
} InnerRange2D;

/* Device array handle (replaces Kokkos::View<flt_t*>, common.h:24-27).
 * `label` may be NULL to skip the view-order check. */
typedef struct view {
    flt_t* data;
    int64_t extent;
    const char* label;
} view;
typedef view mask_view;

#ifdef __cplusplus
/* limiters.h:8-12, test_cases.h:8-14, common.h:30-41 */
enum class Limiter : int { None = 0, Minmod = 1, Superbee = 2 };
enum class Test : int { Sod = 0, Sod_y = 1, Sod_circ = 2, Bizarrium = 3, Sedov = 4 };
enum class Axis : int { X = 0, Y = 1 };
enum class Side : int { Left = 0, Right = 1, Top = 2, Bottom = 3 };
#else
typedef int Limiter;
typedef int Test;
enum { Limiter_None = 0, Limiter_Minmod = 1, Limiter_Superbee = 2 };
enum { Test_Sod = 0, Test_Sod_y = 1, Test_Sod_circ = 2, Test_Bizarrium = 3, Test_Sedov = 4 };
#endif

/* Replaces kernels.h:10-13 (body acoustic.cpp:22-37): first-order acoustic Godunov flux at the faces of `range` */
ARMON_EXPORT void acoustic(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s,
                           ARMON_REF(view) ustar, ARMON_REF(view) pstar,
                           ARMON_CREF(view) rho, ARMON_CREF(view) u, ARMON_CREF(view) pmat, ARMON_CREF(view) cmat);

/* Replaces kernels.h:15-19 (body acoustic.cpp:40-108): second-order GAD flux, None/Minmod/Superbee limiter */
ARMON_EXPORT void acoustic_GAD(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s, flt_t dt, flt_t dx,
                               ARMON_REF(view) ustar, ARMON_REF(view) pstar,
                               ARMON_CREF(view) rho, ARMON_CREF(view) u, ARMON_CREF(view) pmat, ARMON_CREF(view) cmat,
                               Limiter limiter);

/* Replaces kernels.h:21-24 (body fluid_eos.cpp:9-23): perfect gas EOS */
ARMON_EXPORT void perfect_gas_EOS(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, flt_t gamma,
                                  ARMON_CREF(view) rho, ARMON_CREF(view) Emat, ARMON_CREF(view) umat, ARMON_CREF(view) vmat,
                                  ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat);

/* Replaces kernels.h:26-29 (body fluid_eos.cpp:26-62): Bizarrium analytic EOS */
ARMON_EXPORT void bizarrium_EOS(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range,
                                ARMON_CREF(view) rho, ARMON_CREF(view) umat, ARMON_CREF(view) vmat, ARMON_CREF(view) Emat,
                                ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat);

/* Replaces kernels.h:31-33 (body euler_projection.cpp:9-22): Lagrangian cell update, in place */
ARMON_EXPORT void cell_update(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s, flt_t dx, flt_t dt,
                              ARMON_CREF(view) ustar, ARMON_CREF(view) pstar,
                              ARMON_REF(view) rho, ARMON_REF(view) u, ARMON_REF(view) Emat);

/* Replaces kernels.h:35-39 (body euler_projection.cpp:25-48): remap of the conservative variables, in place */
ARMON_EXPORT void euler_projection(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s, flt_t dx, flt_t dt,
                                   ARMON_CREF(view) ustar, ARMON_REF(view) rho, ARMON_REF(view) umat,
                                   ARMON_REF(view) vmat, ARMON_REF(view) Emat,
                                   ARMON_CREF(view) advection_rho, ARMON_CREF(view) advection_urho,
                                   ARMON_CREF(view) advection_vrho, ARMON_CREF(view) advection_Erho);

/* Replaces kernels.h:41-44 (body euler_projection.cpp:51-70): first-order (donor cell) advection fluxes */
ARMON_EXPORT void advection_first_order(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s, flt_t dt,
                                        ARMON_CREF(view) ustar, ARMON_CREF(view) rho, ARMON_CREF(view) umat,
                                        ARMON_CREF(view) vmat, ARMON_CREF(view) Emat,
                                        ARMON_REF(view) advection_rho, ARMON_REF(view) advection_urho,
                                        ARMON_REF(view) advection_vrho, ARMON_REF(view) advection_Erho);

/* Replaces kernels.h:46-49 (body euler_projection.cpp:73-120): second-order (minmod slope) advection fluxes */
ARMON_EXPORT void advection_second_order(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range, Idx s,
                                         flt_t dx, flt_t dt,
                                         ARMON_CREF(view) ustar, ARMON_CREF(view) rho, ARMON_CREF(view) umat,
                                         ARMON_CREF(view) vmat, ARMON_CREF(view) Emat,
                                         ARMON_REF(view) advection_rho, ARMON_REF(view) advection_urho,
                                         ARMON_REF(view) advection_vrho, ARMON_REF(view) advection_Erho);

/* Replaces kernels.h:51-55 (body halo_exchange.cpp:9-36): mirror ghost fill of `stencil_width` layers */
ARMON_EXPORT void boundary_conditions(ARMON_CREF(Range) range, ARMON_CREF(InnerRange1D) inner_range,
                                      Idx stencil_width, Idx disp, flt_t u_factor, flt_t v_factor,
                                      ARMON_REF(view) rho, ARMON_REF(view) umat, ARMON_REF(view) vmat,
                                      ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat,
                                      ARMON_REF(view) Emat);

/* Replaces kernels.h:57-62 (body halo_exchange.cpp:39-64): pack 7 fields x nghost layers of a border */
ARMON_EXPORT void read_border_array(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range,
                                    Idx side_length, Idx nghost,
                                    ARMON_CREF(view) rho, ARMON_CREF(view) umat, ARMON_CREF(view) vmat,
                                    ARMON_CREF(view) pmat, ARMON_CREF(view) cmat, ARMON_CREF(view) gmat,
                                    ARMON_CREF(view) Emat, ARMON_REF(view) value_array);

/* Replaces kernels.h:64-69 (body halo_exchange.cpp:67-92): unpack 7 fields x nghost layers into a border */
ARMON_EXPORT void write_border_array(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range,
                                     Idx side_length, Idx nghost,
                                     ARMON_REF(view) rho, ARMON_REF(view) umat, ARMON_REF(view) vmat,
                                     ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat,
                                     ARMON_REF(view) Emat, ARMON_CREF(view) value_array);

/* Replaces kernels.h:71-77 (body init_test.cpp:51-143): test case initial state over `range` */
ARMON_EXPORT void init_test(ARMON_CREF(Range) range, ARMON_CREF(InnerRange2D) inner_range,
                            Idx row_length, Idx nb_ghosts, Idx nx, Idx ny, flt_t sx,
                            flt_t sy, flt_t ox, flt_t oy, Idx pos_x, Idx pos_y, Idx g_nx, Idx g_ny,
                            ARMON_REF(view) x, ARMON_REF(view) y, ARMON_REF(view) rho, ARMON_REF(view) Emat,
                            ARMON_REF(view) umat, ARMON_REF(view) vmat, ARMON_REF(mask_view) domain_mask,
                            ARMON_REF(view) pmat, ARMON_REF(view) cmat, ARMON_REF(view) gmat,
                            ARMON_REF(view) ustar, ARMON_REF(view) pstar,
                            Test test, bool debug_indexes, flt_t test_option);

/* Replaces kernels.h:79-81 (body reductions.cpp:8-24): CFL time step (min reduction), returned to the host */
ARMON_EXPORT flt_t dt_CFL(ARMON_CREF(Range) range, ARMON_CREF(InnerRange1D) inner_range, flt_t dx, flt_t dy,
                          ARMON_CREF(view) umat, ARMON_CREF(view) vmat, ARMON_CREF(view) cmat,
                          ARMON_CREF(mask_view) domain_mask);

/* Replaces kernels.h:83-86 (body reductions.cpp:27-49): total mass and energy (compensated sums) */
ARMON_EXPORT void conservation_vars(ARMON_CREF(Range) range, ARMON_CREF(InnerRange1D) inner_range, flt_t dx,
                                    ARMON_CREF(view) rho, ARMON_CREF(view) Emat, ARMON_CREF(mask_view) domain_mask,
                                    ARMON_REF(flt_t) total_mass, ARMON_REF(flt_t) total_energy);

/* Replaces utils.h:99-100 / utils.cpp:16-17: error funnel. The default prints and exits(1). A
 * replacement may return: the entry point then returns without launching. */
ARMON_EXPORT extern void (*raise_exception_handler)(const char* kernel, const char* msg);

/* Replace utils.cpp:20-27: ABI queries */
ARMON_EXPORT int idx_size(void);
ARMON_EXPORT bool is_uidx_signed(void);
ARMON_EXPORT int flt_size(void);

#ifdef __cplusplus
}
#endif

#endif /* ARMON_HIP_KERNELS_H */
