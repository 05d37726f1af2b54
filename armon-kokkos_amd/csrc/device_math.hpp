// Per-cell and per-face arithmetic of the Armon Lagrange + remap step, shared by the
// per-stage drop-in kernels and the fused sweep kernels.
//
// Every expression keeps the operation order of the reference kernels (cited per
// function) and the library is compiled with -ffp-contract=off, so that FP64 results
// are bit-identical to an IEEE-strict CPU evaluation of the reference algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define ARMON_HD __host__ __device__ __forceinline__

namespace armon {

// Kokkos::min / Kokkos::max on arithmetic types: (b < a ? b : a) and (a < b ? b : a).
// These fix the sign of zero and the NaN behaviour, unlike fmin/fmax.
template <class T> ARMON_HD T kmin(T a, T b) { return (b < a) ? b : a; }
template <class T> ARMON_HD T kmax(T a, T b) { return (a < b) ? b : a; }

// Correctly rounded y^3 and y^4 (Kokkos::pow(y, 3|4) is the correctly rounded
// libm pow in the reference's IEEE build): double-double products made exact with
// fma, rounded once.
ARMON_HD double cr_pow3(double y)
{
    double p = y * y;
    double e1 = __builtin_fma(y, y, -p);
    double q = p * y;
    double e2 = __builtin_fma(p, y, -q);
    double t = e1 * y;
    return q + (e2 + t);
}

ARMON_HD double cr_pow4(double y)
{
    double p = y * y;
    double e1 = __builtin_fma(y, y, -p);
    double q = p * p;
    double e2 = __builtin_fma(p, p, -q);
    double t = (2.0 * p) * e1;
    return q + (e2 + t);
}

// Division. FP64 (and FP32 outside the FP32-mixed kernels) is the IEEE quotient a / b. The
// FP32-mixed translation unit, which is held to an accuracy bound rather than bit parity, uses the
// native reciprocal instead (about 1 ulp).
template <class T> ARMON_HD T dv(T a, T b) { return a / b; }
#if defined(ARMON_FAST_FP32) && defined(__HIP_DEVICE_COMPILE__)
template <> ARMON_HD float dv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#endif

// x / d from inv = RN(1/d), computed once per cell size: the quotient q = RN(x * inv) and one
// correction with the exact remainder, RN(q - RN(d * q - x) * inv). This is the correctly
// rounded quotient RN(x / d) for every x that recip_div_ok accepts and every d in
// [2^-100, 2^100] (Markstein's correction theorem: inv is within half an ulp of 1/d; when
// |x| >= |d| the first quotient is faithful, and when |x| < |d| it is within 1.5 ulp and the
// correction still cannot cross the midpoint nearest to x / d, since a quotient of two
// 53-bit numbers lies at least 2^-106 / d from any midpoint; the remainder is exact or, in the
// few corner cases near |x| = |d| = 2^k, within the same bound). A zero x gives the signed zero
// of x / d. The caller checks x with recip_div_ok and divides where it fails.
ARMON_HD double div_by_recip(double x, double d, double inv)
{
    const double q = x * inv;
    const double r = __builtin_fma(d, q, -x);
    return __builtin_fma(-r, inv, q);
}

// x is zero or its exponent lies in [-900, 900]: for d in [2^-100, 2^100] the remainder of
// div_by_recip cannot underflow and the quotient cannot overflow.
ARMON_HD bool recip_div_ok(double x)
{
    const uint64_t bits = (uint64_t) __builtin_bit_cast(long long, x);
    const uint32_t e = (uint32_t) (bits >> 52) & 0x7FFu;
    return x == 0 || e - (1023u - 900u) <= 1800u;
}

// Quotients over one divisor (FP64). The compiler divides a by b as
//   b', a' = v_div_scale(b), v_div_scale(a);  y = rcp(b');
//   y = fma(y, fma(-b', y, 1), y), twice;  q = a' * y;  q = v_div_fmas(fma(-b', q, a'), y, q);
//   result = v_div_fixup(q, b, a).
// For b in [2^-200, 2^200] and a zero or of magnitude in [2^-600, 2^300], v_div_scale returns
// its operand unchanged and clears the flag (it acts only on a zero operand, a subnormal b, 1/b
// or a/b, a biased exponent of a at or below 53, or exponents of a and b 768 or more apart),
// v_div_fmas is then the plain FMA, and v_div_fixup returns the quotient with the sign of
// a / b (it substitutes a value only for NaN, infinite or zero operands and for exponents of a
// and b more than 1075 apart). Within those ranges the refined reciprocal y depends on b alone,
// so quotients that share b compute it once and each costs a product and two FMAs, with the
// bits of a / b. The remainder is taken as RN(b q - a) and negated in the last FMA: equal to
// the compiler's RN(a - b q) for a != 0, and for a = +-0 (with b > 0) it makes the result the
// signed zero of a / b. Any operand outside the ranges sends every quotient of the group back
// to a / b. The rcp instruction runs at a quarter of the FMA rate, so a shared divisor saves
// about 9 FMA issue slots per extra quotient.
#if defined(__HIP_DEVICE_COMPILE__)
ARMON_HD double refined_recip(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}

ARMON_HD double refined_quot(double a, double b, double y)
{
    const double q = a * y;
    const double r = __builtin_fma(b, q, -a);
    return __builtin_fma(-r, y, q);
}
#endif

ARMON_HD int shared_divisor_ok(double b) { return (int) (b >= 0x1p-200) & (int) (b <= 0x1p200); }

ARMON_HD int shared_numerator_ok(double a)
{
    const double m = __builtin_fabs(a);
    return ((int) (m >= 0x1p-600) | (int) (a == 0)) & (int) (m <= 0x1p300);
}

// Runs body(div) where div(a) returns a / b for every quotient over b that the body takes.
// The body may run twice (once more with IEEE division when an operand leaves the ranges
// above), so it only assigns its results.
template <class Body>
ARMON_HD void with_divisor(double b, Body&& body)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = refined_recip(b);
    int ok = shared_divisor_ok(b);
    body([&](double a) {
        ok &= shared_numerator_ok(a);
        return refined_quot(a, b, y);
    });
    // marked unlikely: the compiler then keeps this block behind a branch that waves whose lanes
    // are all inside the ranges skip, rather than running both sequences for every lane and
    // selecting
    if (__builtin_expect(!ok, 0)) body([&](double a) { return a / b; });
#else
    body([&](double a) { return a / b; });
#endif
}

// with_divisor for FP64 when SHARED; otherwise (and for the FP32 types) each quotient is
// divided with dv.
template <class T, bool SHARED, class Body>
ARMON_HD void with_divisor_t(T b, Body&& body)
{
    if constexpr (sizeof(T) == 8 && SHARED)
        with_divisor((double) b, static_cast<Body&&>(body));
    else
        body([&](T a) { return dv<T>(a, b); });
}

// The compiler's correctly rounded FP64 square root on gfx950 is rsq followed by two corrections
// (sqrt_core's 10 operations), wrapped by a scaling (operands below 2^-767 are multiplied by
// 2^256 and the root by 2^-128) and a final select that returns the operand itself for +-0 and
// +inf: 21 instructions. For a finite operand of at least 2^-767 the wrappers leave every value
// unchanged, so sqrt_t runs the 10 operations alone there, with the bits of the library square
// root, and the library square root for every other operand (zero, subnormal, tiny, infinite,
// negative or NaN), out of line.
#if defined(__HIP_DEVICE_COMPILE__)
ARMON_HD double sqrt_core(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
#endif

ARMON_HD int sqrt_core_ok(double x) { return (int) (x >= 0x1p-767) & (int) (x < __builtin_huge_val()); }

template <class T> ARMON_HD T sqrt_t(T x);
template <> ARMON_HD double sqrt_t<double>(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_expect(sqrt_core_ok(x), 1)) return sqrt_core(x);
#endif
    return __builtin_sqrt(x);
}
template <> ARMON_HD float sqrt_t<float>(float x) { return __builtin_sqrtf(x); }

template <class T> ARMON_HD T copysign_t(T a, T b);
template <> ARMON_HD double copysign_t<double>(double a, double b) { return __builtin_copysign(a, b); }
template <> ARMON_HD float copysign_t<float>(float a, float b) { return __builtin_copysignf(a, b); }

template <class T> ARMON_HD T abs_t(T a);
template <> ARMON_HD double abs_t<double>(double a) { return __builtin_fabs(a); }
template <> ARMON_HD float abs_t<float>(float a) { return __builtin_fabsf(a); }

enum LimiterKind : int { LIM_NONE = 0, LIM_MINMOD = 1, LIM_SUPERBEE = 2 };

// limiters.h:15-34
template <int L, class T> ARMON_HD T limiter(T r)
{
    if constexpr (L == LIM_MINMOD) {
        return kmax(T(0), kmin(T(1), r));
    } else if constexpr (L == LIM_SUPERBEE) {
        return kmax(kmax(T(0), kmin(T(1), 2 * r)), kmin(T(2), r));
    } else {
        (void) r;
        return T(1);
    }
}

// acoustic.cpp:10-19: linearised acoustic Riemann solver at the face between
// the left cell (…_im) and the right cell (…_i). SHARED: the two quotients share the divisor's
// refined reciprocal (with_divisor), else each is divided.
template <class T, bool SHARED = true>
ARMON_HD void godunov(T rho_i, T rho_im, T c_i, T c_im, T u_i, T u_im, T p_i, T p_im, T& ustar, T& pstar)
{
    T rc_l = rho_im * c_im;
    T rc_r = rho_i * c_i;
    with_divisor_t<T, SHARED>(rc_l + rc_r, [&](auto div) {
        ustar = div(rc_l * u_im + rc_r * u_i + (p_im - p_i));
        pstar = div(rc_r * p_im + rc_l * p_i + rc_l * rc_r * (u_im - u_i));
    });
}

// acoustic.cpp:65-86: GAD second-order correction at face i (between cells i-1 and i),
// from the first-order solutions at faces i-1 (im), i and i+1 (ip).
// Cell values: *_m = cell i-1, *_c = cell i.
template <int L, class T>
ARMON_HD void gad_correct(T dt, T dx,
                          T rho_m, T rho_c, T c_m, T c_c, T u_m, T u_c, T p_m, T p_c,
                          T ustar_im, T pstar_im, T ustar_i, T pstar_i, T ustar_ip, T pstar_ip,
                          T& ustar, T& pstar)
{
    T r_um = dv<T>(ustar_ip - u_c, ustar_i - u_m + T(1e-6));
    T r_pm = dv<T>(pstar_ip - p_c, pstar_i - p_m + T(1e-6));
    T r_up = dv<T>(u_m - ustar_im, u_c - ustar_i + T(1e-6));
    T r_pp = dv<T>(p_m - pstar_im, p_c - pstar_i + T(1e-6));

    r_um = limiter<L>(r_um);
    r_pm = limiter<L>(r_pm);
    r_up = limiter<L>(r_up);
    r_pp = limiter<L>(r_pp);

    T dm_l = rho_m * dx;
    T dm_r = rho_c * dx;
    T Dm = (dm_l + dm_r) / 2;

    T rc_l = rho_m * c_m;
    T rc_r = rho_c * c_c;
    T theta = T(0.5) * (1 - (rc_l + rc_r) / 2 * dv<T>(dt, Dm));

    ustar = ustar_i + theta * (r_up * (u_c - ustar_i) - r_um * (ustar_i - u_m));
    pstar = pstar_i + theta * (r_pp * (p_c - pstar_i) - r_pm * (pstar_i - p_m));
}

// fluid_eos.cpp:16-21
template <class T>
ARMON_HD void perfect_gas(T gamma, T rho, T E, T u, T v, T& p, T& c, T& g)
{
    T e = E - T(0.5) * (u * u + v * v);
    p = (gamma - 1) * rho * e;
    c = sqrt_t<T>(dv<T>(gamma * p, rho));
    g = (1 + gamma) / 2;
}

// fluid_eos.cpp:26-62. Computed in double: `e - eps_k0` and `p - pk0` cancel
// almost completely near the reference density, which FP32 cannot resolve.
ARMON_HD void bizarrium(double rho, double E, double u, double v, double& p_out, double& c_out, double& g_out)
{
    const double rho0 = 1e4, K0 = 1e11, Cv0 = 1e3, T0 = 300, eps0 = 0;
    const double G0 = 1.5, s = 1.5, q = -42080895. / 14941154., r = 727668333. / 149411540.;

    double x = rho / rho0 - 1;
    double g = G0 * (1 - rho0 / rho);

    double f0, f1, f2, f3;
    with_divisor(1 - s * x, [&](auto div) {
        f0 = div(1 + (s / 3 - 2) * x + q * (x * x) + r * (x * x * x));
        f1 = div(s / 3 - 2 + 2 * q * x + 3 * r * (x * x) + s * f0);
        f2 = div(2 * q + 6 * r * x + 2 * s * f1);
        f3 = div(6 * r + 3 * s * f2);
    });

    double eps_k0 = eps0 - Cv0 * T0 * (1 + g) + 0.5 * (K0 / rho0) * (x * x) * f0;
    double pk0 = -Cv0 * T0 * G0 * rho0 + (0.5 * K0 * x * (1 + x) * (1 + x) * (2 * f0 + x * f1));
    double pk0prime = -0.5 * K0 * cr_pow3(1 + x)
            * rho0 * (2 * (1 + 3 * x) * f0 + 2 * x * (2 + 3 * x) * f1 + (x * x) * (1 + x) * f2);
    double xp = x * (1 + x);
    double pk0second = 0.5 * K0 * cr_pow4(1 + x) * (rho0 * rho0)
            * (12 * (1 + 2 * x) * f0 + 6 * (1 + 6 * x + 6 * (x * x)) * f1 + 6 * x * (1 + x) * (1 + 2 * x) * f2
               + (xp * xp) * f3);

    double e = E - 0.5 * (u * u + v * v);
    double p = pk0 + G0 * rho0 * (e - eps_k0);
    double c = sqrt_t<double>(G0 * rho0 * (p - pk0) - pk0prime) / rho;
    double G0rho0 = G0 * rho0;
    g_out = 0.5 / (cr_pow3(rho) * (c * c)) * (pk0second + (G0rho0 * G0rho0) * (p - pk0));
    p_out = p;
    c_out = c;
}

// euler_projection.cpp:17-20: Lagrangian update of a cell from the face fluxes at i and i+1.
template <class T>
ARMON_HD void cell_update(T dx, T dt, T us_i, T us_ip, T ps_i, T ps_ip, T& rho, T& u, T& E)
{
    T dm = rho * dx;
    rho = dv<T>(dm, dx + dt * (us_ip - us_i));
    u += dv<T>(dt, dm) * (ps_i - ps_ip);
    E += dv<T>(dt, dm) * (ps_i * us_i - ps_ip * us_ip);
}

// euler_projection.cpp:73-79
template <class T>
ARMON_HD T slope_minmod(T u_im, T u_i, T u_ip, T r_m, T r_p)
{
    T D_u_p = r_p * (u_ip - u_i);
    T D_u_m = r_m * (u_i - u_im);
    T s = copysign_t<T>(T(1), D_u_p);
    return s * kmax(T(0), kmin(s * D_u_p, s * D_u_m));
}

// euler_projection.cpp:62-69: donor-cell flux, given the donor cell state.
template <class T>
ARMON_HD void advection_flux_1(T disp, T rho, T u, T v, T E, T& f_rho, T& f_urho, T& f_vrho, T& f_Erho)
{
    f_rho = disp * (rho);
    f_urho = disp * (rho * u);
    f_vrho = disp * (rho * v);
    f_Erho = disp * (rho * E);
}

// euler_projection.cpp:91-124: second-order flux at a face. The caller has applied the
// donor shift: cell values *_m, *_c, *_p are cells d-1, d, d+1 around the donor cell d, and
// us_m, us_c, us_p, us_pp are the face velocities at d-1, d, d+1, d+2 (face k sits between
// cells k-1 and k). `Dx` is the signed upwind length chosen by the caller.
template <class T>
ARMON_HD void advection_flux_2(T dx, T dt, T disp, T Dx,
                               T us_m, T us_c, T us_p, T us_pp,
                               T rho_m, T rho_c, T rho_p, T u_m, T u_c, T u_p,
                               T v_m, T v_c, T v_p, T E_m, T E_c, T E_p,
                               T& f_rho, T& f_urho, T& f_vrho, T& f_Erho)
{
    T Dx_lm = dx + dt * (us_c - us_m);
    T Dx_l = dx + dt * (us_p - us_c);
    T Dx_lp = dx + dt * (us_pp - us_p);

    T r_m = dv<T>(2 * Dx_l, Dx_l + Dx_lm);
    T r_p = dv<T>(2 * Dx_l, Dx_l + Dx_lp);

    T slope_r = slope_minmod(rho_m, rho_c, rho_p, r_m, r_p);
    T slope_ur = slope_minmod(rho_m * u_m, rho_c * u_c, rho_p * u_p, r_m, r_p);
    T slope_vr = slope_minmod(rho_m * v_m, rho_c * v_c, rho_p * v_p, r_m, r_p);
    T slope_Er = slope_minmod(rho_m * E_m, rho_c * E_c, rho_p * E_p, r_m, r_p);

    T length_factor = dv<T>(Dx, 2 * Dx_l);
    f_rho = disp * (rho_c - slope_r * length_factor);
    f_urho = disp * (rho_c * u_c - slope_ur * length_factor);
    f_vrho = disp * (rho_c * v_c - slope_vr * length_factor);
    f_Erho = disp * (rho_c * E_c - slope_Er * length_factor);
}

// How the projection divides by dx: DX_DIVIDE (x / dx), DX_POW2 (dx is a power of two and
// inv_dx = 1 / dx, so x / dx == x * inv_dx exactly), DX_RECIP (inv_dx = RN(1 / dx) with dx in
// [2^-100, 2^100]: div_by_recip, the correctly rounded quotient, for the numerators that
// recip_div_ok accepts; FP64 only).
enum DxMode : int { DX_DIVIDE = 0, DX_POW2 = 1, DX_RECIP = 2 };

// euler_projection.cpp:42-52: remap of cell i from the advection fluxes at faces i, i+1.
// SHARED: u, v and E share the new density's refined reciprocal (with_divisor).
template <int DXM, class T, bool SHARED = true>
ARMON_HD void projection(T dx, T inv_dx, T dt, T us_i, T us_ip,
                         T f_rho_i, T f_rho_ip, T f_urho_i, T f_urho_ip,
                         T f_vrho_i, T f_vrho_ip, T f_Erho_i, T f_Erho_ip,
                         T& rho, T& u, T& v, T& E)
{
    T dX = dx + dt * (us_ip - us_i);
    T n_rho, n_urho, n_vrho, n_Erho;
    if constexpr (DXM == DX_POW2) {
        n_rho = (dX * rho - (f_rho_ip - f_rho_i)) * inv_dx;
        n_urho = (dX * rho * u - (f_urho_ip - f_urho_i)) * inv_dx;
        n_vrho = (dX * rho * v - (f_vrho_ip - f_vrho_i)) * inv_dx;
        n_Erho = (dX * rho * E - (f_Erho_ip - f_Erho_i)) * inv_dx;
    } else if constexpr (DXM == DX_RECIP && sizeof(T) == 8) {
        const T a_rho = dX * rho - (f_rho_ip - f_rho_i);
        const T a_urho = dX * rho * u - (f_urho_ip - f_urho_i);
        const T a_vrho = dX * rho * v - (f_vrho_ip - f_vrho_i);
        const T a_Erho = dX * rho * E - (f_Erho_ip - f_Erho_i);
        if ((int) recip_div_ok(a_rho) & (int) recip_div_ok(a_urho) & (int) recip_div_ok(a_vrho) &
            (int) recip_div_ok(a_Erho)) {
            n_rho = div_by_recip(a_rho, dx, inv_dx);
            n_urho = div_by_recip(a_urho, dx, inv_dx);
            n_vrho = div_by_recip(a_vrho, dx, inv_dx);
            n_Erho = div_by_recip(a_Erho, dx, inv_dx);
        } else {
            n_rho = dv<T>(a_rho, dx);
            n_urho = dv<T>(a_urho, dx);
            n_vrho = dv<T>(a_vrho, dx);
            n_Erho = dv<T>(a_Erho, dx);
        }
    } else {
        (void) inv_dx;
        n_rho = dv<T>(dX * rho - (f_rho_ip - f_rho_i), dx);
        n_urho = dv<T>(dX * rho * u - (f_urho_ip - f_urho_i), dx);
        n_vrho = dv<T>(dX * rho * v - (f_vrho_ip - f_vrho_i), dx);
        n_Erho = dv<T>(dX * rho * E - (f_Erho_ip - f_Erho_i), dx);
    }
    rho = n_rho;
    with_divisor_t<T, SHARED>(n_rho, [&](auto div) {
        u = div(n_urho);
        v = div(n_vrho);
        E = div(n_Erho);
    });
}

// reductions.cpp:18-20: CFL bound of one cell (mask 1 for real cells).
template <class T>
ARMON_HD T cfl_cell(T dx, T dy, T u, T v, T c, T mask)
{
    T max_cx = kmax(abs_t<T>(u + c), abs_t<T>(u - c)) * mask;
    T max_cy = kmax(abs_t<T>(v + c), abs_t<T>(v - c)) * mask;
    return kmin(dv<T>(dx, max_cx), dv<T>(dy, max_cy));
}

}  // namespace armon
