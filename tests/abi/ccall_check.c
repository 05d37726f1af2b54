/*
 * The Julia calling convention of Armon.jl (ccall through Kokkos.jl, reference README.md:5-7),
 * exercised from C: the library is opened with dlopen and every entry point resolved with dlsym,
 * with no header; Range / InnerRange / view are plain structs passed by pointer (Julia's
 * Ref{T}); the error funnel is replaced by storing a function pointer into the data symbol
 * raise_exception_handler (Julia: unsafe_store!(cglobal(...), @cfunction(...)), reference
 * utils.cpp:16-17).
 *
 * Test infrastructure: run by tests/test_gpu_abi.py. It checks, against the oracle library
 * (opened the same way):
 *   1. the ABI queries of utils.cpp:20-27;
 *   2. acoustic_GAD (both axes) and dt_CFL on seeded random states, bit for bit;
 *   3. an invalid limiter reaches the replaced handler with the reference's message
 *      (acoustic.cpp:101-106), both with a handler that returns (the entry point then returns
 *      without launching) and with one that does not return (longjmp, as a Julia exception
 *      unwinds), after which the library keeps working.
 * Usage: ccall_check <libarmon.so> <liboracle.so>. Prints one line per check; exit status 0 when
 * every check passes.
 */
#include <dlfcn.h>
#include <setjmp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int64_t start, end; } Range;
typedef struct { int64_t start, step; } InnerRange1D;
typedef struct { int64_t main_range_start, main_range_step, row_range_start; uint64_t row_range_length; } InnerRange2D;
typedef struct { double* data; int64_t extent; const char* label; } View;

typedef void (*handler_t)(const char*, const char*);
typedef void (*acoustic_GAD_t)(const Range*, const InnerRange2D*, int64_t, double, double, View*, View*,
                               const View*, const View*, const View*, const View*, int);
typedef double (*dt_CFL_t)(const Range*, const InnerRange1D*, double, double, const View*, const View*,
                           const View*, const View*);
typedef int (*int_query_t)(void);
typedef _Bool (*bool_query_t)(void);
typedef void* (*malloc_t)(size_t);
typedef int (*free_t)(void*);
typedef int (*copy_t)(void*, const void*, size_t);
typedef int (*sync_t)(void);
typedef int (*o_gad_t)(const Range*, const InnerRange2D*, int64_t, double, double, double*, double*,
                       const double*, const double*, const double*, const double*, int);
typedef double (*o_dt_t)(const Range*, const InnerRange1D*, double, double, const double*, const double*,
                         const double*, const double*);

static int failures = 0;
#define CHECK(cond, ...)                                       \
    do {                                                       \
        if (cond) printf("ok   ");                             \
        else { printf("FAIL "); failures++; }                  \
        printf(__VA_ARGS__);                                   \
        printf("\n");                                          \
        fflush(stdout);                                        \
    } while (0)

static void* sym(void* lib, const char* name)
{
    void* p = dlsym(lib, name);
    if (!p) {
        printf("FAIL missing symbol %s\n", name);
        exit(2);
    }
    return p;
}

/* ---- handlers ------------------------------------------------------------------------- */
static char last_kernel[128], last_msg[512];
static int handler_calls = 0;
static jmp_buf unwind;

static void returning_handler(const char* kernel, const char* msg)
{
    handler_calls++;
    snprintf(last_kernel, sizeof last_kernel, "%s", kernel);
    snprintf(last_msg, sizeof last_msg, "%s", msg);
}

static void unwinding_handler(const char* kernel, const char* msg)
{
    returning_handler(kernel, msg);
    longjmp(unwind, 1);
}

/* ---- data --------------------------------------------------------------------------------- */
static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double uniform(double lo, double hi)
{
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return lo + (hi - lo) * (double) (rng >> 11) / 9007199254740992.0;
}

enum { NX = 300, NY = 200, G = 5, RL = NX + 2 * G, CL = NY + 2 * G, N = RL * CL };

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s <libarmon.so> <liboracle.so>\n", argv[0]);
        return 2;
    }
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    void* olib = dlopen(argv[2], RTLD_NOW | RTLD_LOCAL);
    if (!lib || !olib) {
        printf("FAIL dlopen: %s\n", dlerror());
        return 2;
    }
    /* 1. ABI queries (utils.cpp:20-27) */
    CHECK(((int_query_t) sym(lib, "idx_size"))() == 8, "idx_size() == 8");
    CHECK(((int_query_t) sym(lib, "flt_size"))() == 8, "flt_size() == 8");
    CHECK(((bool_query_t) sym(lib, "is_uidx_signed"))() == 0, "is_uidx_signed() == false");

    /* the error funnel: a data symbol holding a function pointer */
    handler_t* handler = (handler_t*) sym(lib, "raise_exception_handler");
    CHECK(*handler != NULL, "raise_exception_handler holds the default handler");
    *handler = returning_handler;

    acoustic_GAD_t acoustic_GAD = (acoustic_GAD_t) sym(lib, "acoustic_GAD");
    dt_CFL_t dt_CFL = (dt_CFL_t) sym(lib, "dt_CFL");
    malloc_t dmalloc = (malloc_t) sym(lib, "armon_malloc");
    free_t dfree = (free_t) sym(lib, "armon_free");
    copy_t htod = (copy_t) sym(lib, "armon_memcpy_htod");
    copy_t dtoh = (copy_t) sym(lib, "armon_memcpy_dtoh");
    sync_t dsync = (sync_t) sym(lib, "armon_synchronize");
    o_gad_t o_acoustic_GAD = (o_gad_t) sym(olib, "o_acoustic_GAD");
    o_dt_t o_dt_CFL = (o_dt_t) sym(olib, "o_dt_CFL");

    /* host arrays and their device copies, labelled as the reference labels its views */
    const char* names[8] = {"ustar", "pstar", "rho", "u", "pmat", "cmat", "vmat", "domain_mask"};
    const double lo[8] = {-1, 0.1, 0.5, -1, 0.1, 0.5, -1, 0};
    const double hi[8] = {1, 2, 2, 1, 2, 2, 1, 0};
    static double host[8][N], back[N];
    View v[8];
    char labels[8][32];
    for (int f = 0; f < 8; f++) {
        for (int k = 0; k < N; k++) host[f][k] = uniform(lo[f], hi[f]);
        snprintf(labels[f], sizeof labels[f], "Armon_%s", names[f]);
        v[f].data = (double*) dmalloc(sizeof(double) * N);
        v[f].extent = N;
        v[f].label = labels[f];
        if (!v[f].data || htod(v[f].data, host[f], sizeof(double) * N)) {
            printf("FAIL device allocation or copy\n");
            return 2;
        }
    }
    for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) host[7][(j + G) * RL + i + G] = 1.0;   /* real cells */
    htod(v[7].data, host[7], sizeof(double) * N);

    /* 2. acoustic_GAD over the flux range of each axis (ranges.cpp:14-30 with Euler 2nd order:
     * the real domain inflated by 2 and expanded by 1 along the axis), superbee; the oracle's
     * ustar and pstar go through the same two calls */
    static double o_us[N], o_ps[N];
    memcpy(o_us, host[0], sizeof o_us);
    memcpy(o_ps, host[1], sizeof o_ps);
    for (int axis = 0; axis < 2; axis++) {
        const int64_t s = axis == 0 ? 1 : RL;
        const int64_t first = (int64_t) G * RL + G;   /* index of real cell (0, 0) */
        InnerRange2D ir;
        Range r;
        if (axis == 0) {
            ir = (InnerRange2D){first, RL, -2, NX + 5};
            r = (Range){0, (int64_t) NY * (NX + 5)};
        } else {
            ir = (InnerRange2D){first - 2 * RL, RL, 0, NX};
            r = (Range){0, (int64_t) (NY + 5) * NX};
        }
        const double dt = 0.0123, dx = axis == 0 ? 1.0 / NX : 1.0 / NY;
        acoustic_GAD(&r, &ir, s, dt, dx, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], 2);
        const int orc = o_acoustic_GAD(&r, &ir, s, dt, dx, o_us, o_ps, host[2], host[3], host[4], host[5], 2);
        dsync();
        int same = orc == 0;
        dtoh(back, v[0].data, sizeof back);
        same = same && memcmp(back, o_us, sizeof back) == 0;
        dtoh(back, v[1].data, sizeof back);
        same = same && memcmp(back, o_ps, sizeof back) == 0;
        CHECK(same && handler_calls == 0, "acoustic_GAD axis %d through dlsym: ustar, pstar bit-identical to the oracle",
              axis);
    }

    /* dt_CFL over the real domain's 1D iteration (indexing.cpp:83-89); the x velocity is the
     * array passed as `u` above, under the label dt_CFL checks */
    View umat = v[3];
    umat.label = "Armon_umat";
    {
        const Range r = {0, (int64_t) (NY - 1) * RL + NX};
        const InnerRange1D ir = {(int64_t) G * RL + G, 1};
        handler_calls = 0;
        const double got = dt_CFL(&r, &ir, 1.0 / NX, 1.0 / NY, &umat, &v[6], &v[5], &v[7]);
        const double want = o_dt_CFL(&r, &ir, 1.0 / NX, 1.0 / NY, host[3], host[6], host[5], host[7]);
        CHECK(got == want && got > 0 && handler_calls == 0, "dt_CFL through dlsym: %.17g == oracle %.17g", got,
              want);
    }

    /* 3. an invalid limiter reaches the replaced handler */
    const Range r1 = {0, 10};
    const InnerRange2D ir1 = {(int64_t) G * RL + G, RL, 0, 10};
    memcpy(back, host[0], sizeof back);
    htod(v[0].data, back, sizeof back);
    handler_calls = 0;
    acoustic_GAD(&r1, &ir1, 1, 0.01, 0.01, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], 7);
    CHECK(handler_calls == 1 && strcmp(last_kernel, "acoustic_GAD") == 0 &&
                  strstr(last_msg, "Invalid limiter index: 7") != NULL,
          "invalid limiter -> returning handler: (%s) %s", last_kernel, last_msg);
    dsync();
    dtoh(back, v[0].data, sizeof back);
    CHECK(memcmp(back, host[0], sizeof back) == 0, "no launch after the handler returned (ustar unchanged)");

    *handler = unwinding_handler;
    handler_calls = 0;
    if (setjmp(unwind) == 0) {
        acoustic_GAD(&r1, &ir1, 1, 0.01, 0.01, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], -1);
        CHECK(0, "the unwinding handler did not run");
    } else {
        CHECK(handler_calls == 1 && strstr(last_msg, "Invalid limiter index: -1") != NULL,
              "invalid limiter -> unwinding handler: (%s) %s", last_kernel, last_msg);
    }
    /* a view passed in the wrong order: the label check (utils.h:80-93) */
    handler_calls = 0;
    if (setjmp(unwind) == 0) {
        acoustic_GAD(&r1, &ir1, 1, 0.01, 0.01, &v[1], &v[0], &v[2], &v[3], &v[4], &v[5], 1);
        CHECK(0, "the wrong view order was not reported");
    } else {
        CHECK(handler_calls == 1 && strstr(last_msg, "Wrong view order") != NULL, "wrong view order -> handler: %s",
              last_msg);
    }
    /* the library keeps working after a handler unwound out of it */
    *handler = returning_handler;
    {
        const Range r = {0, (int64_t) (NY - 1) * RL + NX};
        const InnerRange1D ir = {(int64_t) G * RL + G, 1};
        handler_calls = 0;
        const double got = dt_CFL(&r, &ir, 1.0 / NX, 1.0 / NY, &umat, &v[6], &v[5], &v[7]);
        const double want = o_dt_CFL(&r, &ir, 1.0 / NX, 1.0 / NY, host[3], host[6], host[5], host[7]);
        CHECK(got == want && handler_calls == 0, "dt_CFL after the unwinding: %.17g == oracle %.17g", got, want);
    }
    for (int f = 0; f < 8; f++) dfree(v[f].data);
    printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
