// Internal helpers of the Armon HIP library: error funnel, library stream, device scratch,
// and the host-side index-span checks that turn out-of-bounds accesses into errors
// before any kernel is launched.
#pragma once

#include "armon/kernels.h"
#include "armon/armon.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace armon {

// Error classes reported through raise_exception_handler as "(type) message",
// like the reference's KERNEL_TRY/KERNEL_CATCH (utils.h:103-114, utils.cpp:49-54).
struct Error : std::runtime_error {
    std::string type;
    Error(const std::string& t, const std::string& msg) : std::runtime_error(msg), type(t) {}
};

inline Error invalid_argument(const std::string& m) { return Error("std::invalid_argument", m); }
inline Error out_of_range(const std::string& m) { return Error("std::out_of_range", m); }
inline Error hip_failure(const std::string& m) { return Error("hip_error", m); }

// Reports `e` through the handler; also records it for armon_last_error().
void report(const char* kernel, const Error& e);
void report_std(const char* kernel, const std::exception& e);
// Errors reported so far in this process (a caller compares it before and after a call).
uint64_t report_count();

inline void hip_check(hipError_t err, const char* what)
{
    if (err != hipSuccess) throw hip_failure(std::string(what) + ": " + hipGetErrorString(err));
}

hipStream_t stream();

// The boundary velocity factors of `side` (ARMON_SIDE_*): the caller's when p.bc_override is
// set, else the test's (armon_test_boundary_condition).
inline void boundary_factors(const ArmonParams& p, int side, double* uf, double* vf)
{
    if (p.bc_override) {
        *uf = p.bc_factors[side][0];
        *vf = p.bc_factors[side][1];
    } else {
        armon_test_boundary_condition(p.test, side, uf, vf);
    }
}

// Library-owned device scratch (reduction partials and scalars), one per device,
// allocated on first use and kept until the process ends.
struct Scratch {
    double* partials = nullptr;       // reduction partials, kPartials entries
    unsigned long long* min_bits = nullptr;  // min-reduction accumulator (bit pattern of a non-negative double)
    double* result = nullptr;         // final reduction results (up to 8 doubles)
    static constexpr int kPartials = 4 * 1024;
};
Scratch& scratch();

// Inclusive index span [lo, hi] of a set of array indices; empty if lo > hi.
struct Span {
    int64_t lo = 1, hi = 0;
    bool empty() const { return lo > hi; }
    void add(int64_t i)
    {
        if (empty()) { lo = hi = i; }
        else { lo = i < lo ? i : lo; hi = i > hi ? i : hi; }
    }
    Span shifted(int64_t off_lo, int64_t off_hi) const
    {
        Span s = *this;
        if (!empty()) { s.lo += off_lo; s.hi += off_hi; }
        return s;
    }
    Span merged(const Span& o) const
    {
        Span s = *this;
        if (!o.empty()) { s.add(o.lo); s.add(o.hi); }
        return s;
    }
};

// Launch geometry of a Range scaled through an InnerRange2D (indexing.h:67-72):
// linear k in [start, end) maps to row = k / len, col = k % len,
// index = base + row * step + col with base = row_range_start + main_range_start.
struct Iter2D {
    int64_t start, end, len, first_row, last_row, base, step;
};

struct Iter1D {
    int64_t start, end, step, base;
};

Iter2D make_iter2d(const Range& r, const InnerRange2D& ir);
Iter1D make_iter1d(const Range& r, const InnerRange1D& ir);
Span span_of(const Iter2D& it);
Span span_of(const Iter1D& it);

// Checks that every index of `s` is inside the view, and (when both labels are
// set) that the view label ends with `name` (utils.h:80-93).
void check_view(const view& v, const char* name, const Span& s, bool check_label = true);

inline int64_t abs64(int64_t v) { return v < 0 ? -v : v; }

}  // namespace armon

#define ARMON_KERNEL_TRY try
#define ARMON_KERNEL_CATCH(name)                                    \
    catch (const armon::Error& err) { armon::report(name, err); }   \
    catch (const std::exception& err) { armon::report_std(name, err); }
