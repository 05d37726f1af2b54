// Command-line driver: the flags of the reference (src/main_2D.cpp:9-38) plus the execution
// options of this implementation (--engine, --arith, --graph-chunk, --y-run).
#include "armon/armon.h"

#include <chrono>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

static const char USAGE[] = R"(
 == Armon (MI355X) ==
2D compressible Euler solver, dimensionally split Lagrange + remap, on AMD gfx950 GPUs.

Options:
    -h or --help            Prints this message and exit
    -t <test>               Test case: 'Sod', 'Sod_y', 'Sod_circ', 'Bizarrium' or 'Sedov'
    -s <scheme>             Numeric scheme: 'Godunov' (first order) or 'GAD' (second order)
    --limiter <limiter>     Limiter for the second order scheme: 'None', 'Minmod' (default), 'Superbee'
    --cells Nx,Ny           Number of cells in the 2D mesh
    --nghost N              Number of ghost cells around the 2D domain
    --stencil N             Width of the stencil in cells. Defaults to the minimum number of ghost cells.
    --cycle N               Maximum number of iterations
    --riemann <solver>      Riemann solver: 'acoustic' only
    --projection <scheme>   Projection scheme: 'euler' (1st order) or 'euler_2nd' (2nd order)
    --time T                Maximum simulation time (in seconds)
    --cfl C                 CFL number
    --dt T                  Initial time step (in seconds)
    --cst-dt 0-1            Constant time step mode
    --write-output 0-1      If the variables should be written to the output file
    --write-ghosts 0-1      Include the ghost cells in the output file
    --output <file>         The output file name/path
    --precision D           Number of digits of precision printed to the output file for each value.
    --write-throughput 0-1  Enable writing the cell throughput to a separate file (in Mega cells/sec)
    --splitting <method>    Axis splitting method: 'Sequential', 'SequentialSym', 'Strang', 'X_only', 'Y_only'
    --compare 0-1           Compare each step of the solver with reference files <output>_<cycle>_<step>_<axis>
                            (staged engine)
    --comparison-tolerance T  Relative tolerance of the comparison (default 1e-13)
    --verbose 0-3           Verbosity (0: high, 3: low)
    --engine <engine>       'fused' (default: one kernel per axis sweep) or 'staged' (one launch per stage)
    --arith <mode>          'fp64' (default) or 'fp32mixed' (fused engine only)
    --graph-chunk N         Cycles per captured hipGraph launch (fused engine; 0 disables)
    --y-run N               Rows per marching run of the fused Y sweep (0: chosen from the domain size)
    --batch N               Run a batched ensemble of N domains (64^2 to 1024^2 cells, Sod / Sod_circ /
                            Sedov, perturbed states) drawn from --seed S (default 2024), each to its own
                            end time or --cycle; --autotune 0-1 (default 1)
)";

// Batched-ensemble mode (--batch N): N domains drawn from --seed, advanced together.
struct BatchOptions {
    int domains = 0;
    unsigned long long seed = 2024;
    int autotune = 1;
};

static bool parse_arguments(ArmonParams& p, BatchOptions& bo, int argc, char** argv)
{
    errno = 0;
    bool engine_given = false;
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        const char* v = i + 1 < argc ? argv[i + 1] : nullptr;
        auto need = [&]() {
            if (!v) fprintf(stderr, "Missing value for option: %s\n", a);
            return v != nullptr;
        };
        if (strcmp(a, "-h") == 0 || strcmp(a, "--help") == 0) {
            puts(USAGE);
            exit(0);
        } else if (strcmp(a, "-t") == 0) {
            if (!need()) return false;
            if (strcmp(v, "Sod") == 0) p.test = 0;
            else if (strcmp(v, "Sod_y") == 0) p.test = 1;
            else if (strcmp(v, "Sod_circ") == 0) p.test = 2;
            else if (strcmp(v, "Bizarrium") == 0) p.test = 3;
            else if (strcmp(v, "Sedov") == 0) p.test = 4;
            else { printf("Unknown test: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "-s") == 0) {
            if (!need()) return false;
            if (strcmp(v, "Godunov") == 0) p.scheme = ARMON_SCHEME_GODUNOV;
            else if (strcmp(v, "GAD") == 0) p.scheme = ARMON_SCHEME_GAD;
            else { printf("Wrong scheme: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--limiter") == 0) {
            if (!need()) return false;
            if (strcmp(v, "None") == 0) p.limiter = 0;
            else if (strcmp(v, "Minmod") == 0) p.limiter = 1;
            else if (strcmp(v, "Superbee") == 0) p.limiter = 2;
            else { printf("Wrong limiter: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--cells") == 0) {
            if (!need()) return false;
            char* comma = nullptr;
            p.nx = (int) strtol(v, &comma, 10);
            p.ny = (comma && *comma == ',') ? (int) strtol(comma + 1, nullptr, 10) : p.nx;
            i++;
        } else if (strcmp(a, "--nghost") == 0) {
            if (!need()) return false;
            p.nb_ghosts = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--stencil") == 0) {
            if (!need()) return false;
            p.stencil_width = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--cycle") == 0) {
            if (!need()) return false;
            p.max_cycles = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--verbose") == 0) {
            if (!need()) return false;
            p.verbose = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--riemann") == 0) {
            if (!need()) return false;
            if (strcmp(v, "acoustic") != 0) { printf("Wrong Riemann solver: %s\n", v); return false; }
            p.riemann = ARMON_RIEMANN_ACOUSTIC;
            i++;
        } else if (strcmp(a, "--time") == 0) {
            if (!need()) return false;
            p.max_time = strtod(v, nullptr);
            i++;
        } else if (strcmp(a, "--cfl") == 0) {
            if (!need()) return false;
            p.cfl = strtod(v, nullptr);
            i++;
        } else if (strcmp(a, "--dt") == 0) {
            if (!need()) return false;
            p.Dt = strtod(v, nullptr);
            i++;
        } else if (strcmp(a, "--write-output") == 0) {
            if (!need()) return false;
            p.write_output = (int) strtol(v, nullptr, 2);
            i++;
        } else if (strcmp(a, "--write-ghosts") == 0) {
            if (!need()) return false;
            p.write_ghosts = (int) strtol(v, nullptr, 2);
            i++;
        } else if (strcmp(a, "--precision") == 0) {
            if (!need()) return false;
            p.output_precision = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--projection") == 0) {
            if (!need()) return false;
            if (strcmp(v, "none") == 0) p.projection = ARMON_PROJ_NONE;
            else if (strcmp(v, "euler") == 0) p.projection = ARMON_PROJ_EULER;
            else if (strcmp(v, "euler_2nd") == 0) p.projection = ARMON_PROJ_EULER_2ND;
            else { printf("Wrong projection scheme: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--cst-dt") == 0) {
            if (!need()) return false;
            p.cst_dt = (int) strtol(v, nullptr, 2);
            i++;
        } else if (strcmp(a, "--output") == 0) {
            if (!need()) return false;
            p.output_file = v;
            i++;
        } else if (strcmp(a, "--write-throughput") == 0) {
            if (!need()) return false;
            p.write_throughput = (int) strtol(v, nullptr, 2);
            i++;
        } else if (strcmp(a, "--splitting") == 0) {
            if (!need()) return false;
            if (strcmp(v, "Sequential") == 0) p.axis_splitting = ARMON_SPLIT_SEQUENTIAL;
            else if (strcmp(v, "SequentialSym") == 0) p.axis_splitting = ARMON_SPLIT_SEQUENTIAL_SYM;
            else if (strcmp(v, "Strang") == 0) p.axis_splitting = ARMON_SPLIT_STRANG;
            else if (strcmp(v, "X_only") == 0) p.axis_splitting = ARMON_SPLIT_X_ONLY;
            else if (strcmp(v, "Y_only") == 0) p.axis_splitting = ARMON_SPLIT_Y_ONLY;
            else { printf("Wrong axis splitting method: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--compare") == 0) {
            if (!need()) return false;
            p.compare = (int) strtol(v, nullptr, 2);
            i++;
        } else if (strcmp(a, "--comparison-tolerance") == 0) {
            if (!need()) return false;
            p.comparison_tolerance = strtod(v, nullptr);
            i++;
        } else if (strcmp(a, "--engine") == 0) {
            if (!need()) return false;
            engine_given = true;
            if (strcmp(v, "fused") == 0) p.engine = ARMON_ENGINE_FUSED;
            else if (strcmp(v, "staged") == 0) p.engine = ARMON_ENGINE_STAGED;
            else { printf("Wrong engine: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--arith") == 0) {
            if (!need()) return false;
            if (strcmp(v, "fp64") == 0) p.precision = ARMON_PREC_FP64;
            else if (strcmp(v, "fp32mixed") == 0) p.precision = ARMON_PREC_FP32_MIXED;
            else { printf("Wrong arithmetic mode: %s\n", v); return false; }
            i++;
        } else if (strcmp(a, "--graph-chunk") == 0) {
            if (!need()) return false;
            p.graph_chunk = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--y-run") == 0) {
            if (!need()) return false;
            p.y_run_rows = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--batch") == 0) {
            if (!need()) return false;
            bo.domains = (int) strtol(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--seed") == 0) {
            if (!need()) return false;
            bo.seed = strtoull(v, nullptr, 10);
            i++;
        } else if (strcmp(a, "--autotune") == 0) {
            if (!need()) return false;
            bo.autotune = (int) strtol(v, nullptr, 2);
            i++;
        } else {
            fprintf(stderr, "Wrong option: %s\n", a);
            return false;
        }
    }
    if (errno != 0) {
        fprintf(stderr, "Parsing error occurred: %s\n", strerror(errno));
        return false;
    }
    // the per-step comparison checks every stage, and the fused engine keeps no ghost cells in
    // memory: both run the staged engine unless told otherwise
    if ((p.compare || (p.write_output && p.write_ghosts)) && !engine_given) p.engine = ARMON_ENGINE_STAGED;
    return true;
}

// The ensemble: domains of 64^2 to 1024^2 cells from the seed, each run to its own end time
// (or --cycle cycles) with the scheme options of the command line.
static int run_batch(ArmonParams common, const BatchOptions& bo)
{
    const int n = bo.domains;
    ArmonDomainSpec* specs = new ArmonDomainSpec[n];
    int status = 1;
    ArmonBatch* b = nullptr;
    if (armon_batch_random_specs(bo.seed, n, 64 * 64, 1024 * 1024, 4.0, specs) != 0) goto done;
    common.engine = ARMON_ENGINE_FUSED;
    b = armon_batch_create(&common, n, specs);
    if (!b) goto done;
    if (armon_batch_init(b) != 0) goto done;
    if (bo.autotune) {
        ArmonBatchTuning t;
        if (armon_batch_autotune(b, &t) != 0) goto done;
        if (common.verbose < 3)
            printf("Autotuned: X tiles of %d cells at %d per thread, Y runs of %d rows, X grid %d, Y grid %d "
                   "(%.3f + %.3f ms per sweep pair)\n", t.tile_x, t.x_cells_per_thread, t.y_run_rows, t.grid_x,
                   t.grid_y, t.ms_x, t.ms_y);
    }
    {
        if (armon_batch_begin(b) != 0) goto done;
        const auto t0 = std::chrono::steady_clock::now();
        const int run_status = armon_batch_run(b);   // returns once every domain has stopped
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const int finish_status = armon_batch_finish(b);
        if (run_status != 0 || finish_status != 0) goto done;
        double updates = 0;
        int max_cycles = 0;
        for (int d = 0; d < n; d++) {
            ArmonResult r;
            ArmonParams q;
            armon_batch_domain_result(b, d, &r);
            armon_batch_domain_params(b, d, &q);
            updates += (double) r.cycles * q.nx * q.ny;
            if (r.cycles > max_cycles) max_cycles = r.cycles;
            if (common.verbose < 2)
                printf("domain %3d: test %d, %4dx%-4d, %6d cycles, t = %.6g, final dt = %.6g\n", d, q.test, q.nx,
                       q.ny, r.cycles, r.time, r.final_dt);
        }
        if (common.verbose < 4)
            printf("\nBatch of %d domains (%lld cells): %.4g s, up to %d cycles, %.4g Mega cell-updates/sec\n", n,
                   (long long) armon_batch_cells(b), ms * 1e-3, max_cycles, updates / (ms * 1e-3) / 1e6);
        status = 0;
    }
done:
    if (status != 0) fprintf(stderr, "batch run failed: %s\n", armon_last_error());
    if (b) armon_batch_destroy(b);
    delete[] specs;
    return status;
}

int main(int argc, char* argv[])
{
    ArmonParams p;
    armon_params_default(&p);
    BatchOptions bo;
    if (!parse_arguments(p, bo, argc, argv)) return 1;
    if (bo.domains > 0) return run_batch(p, bo);
    if (armon_params_init(&p) != 0) return 1;
    if (!armon_params_check(&p)) return 1;
    if (p.verbose < 3) {
        printf("Parameters:\n");
        printf(" - engine:     %s, %s\n", p.engine == ARMON_ENGINE_FUSED ? "fused" : "staged",
               p.precision == ARMON_PREC_FP64 ? "fp64" : "fp32-mixed");
        printf(" - test:       %d, scheme %s, limiter %d, projection %d, splitting %d\n", p.test,
               p.scheme == ARMON_SCHEME_GAD ? "GAD" : "Godunov", p.limiter, p.projection, p.axis_splitting);
        printf(" - domain:     %dx%d (%d ghosts), size %gx%g, origin (%g, %g)\n", p.nx, p.ny, p.nb_ghosts,
               p.domain_size[0], p.domain_size[1], p.domain_origin[0], p.domain_origin[1]);
        printf(" - CFL:        %g, stencil width %d, max time %g, max cycles %d\n", p.cfl, p.stencil_width,
               p.max_time, p.max_cycles);
    }
    ArmonResult r{};
    const int status = armon_run(&p, &r);
    if (status == 1) printf("Invalid dt at cycle %d\n", r.cycles);
    return status == 0 ? 0 : 1;
}
