// Stage tracing of the host driver.
//
//  - roctx ranges named as the reference's NVTX ranges (armon_2D.cpp:19-54, 306-322, 339, 394-404):
//    cycle, time_step, axis, EOS, BC, fluxes, update, remap, EOS_init, init_test. They cost
//    nothing unless a tool listens (rocprofv3 --marker-trace).
//  - a per-stage time table, the counterpart of the reference's TIC/TAC table
//    (armon_2D.cpp:57-67, printed at :420-428 when verbose < 3), under the reference's labels.
//    The reference times each stage on the host without a fence, which on a GPU measures the
//    launch; here each stage is bracketed by HIP events on the solver stream, so the table holds
//    device time.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <map>
#include <string>
#include <vector>

namespace armon {

class StageClock {
public:
    struct Entry {
        double ms = 0;
        long calls = 0;
    };

    void set_stream(hipStream_t s) { stream_ = s; }
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }

    // Records the start of stage `label` on the stream; returns a token for stop().
    int start(const char* label)
    {
        if (!on_) return -1;
        Pending p{label, take(), take()};
        if (hipEventRecord(p.a, stream_) != hipSuccess) return -1;
        pending_.push_back(p);
        return (int) pending_.size() - 1;
    }

    void stop(int token)
    {
        if (token < 0 || token >= (int) pending_.size()) return;
        (void) hipEventRecord(pending_[token].b, stream_);
    }

    // Adds the durations of the stages whose end event has completed; with `wait`, waits for all.
    void collect(bool wait)
    {
        std::vector<Pending> keep;
        for (Pending& p : pending_) {
            if (wait) (void) hipEventSynchronize(p.b);
            if (hipEventQuery(p.b) != hipSuccess) {
                keep.push_back(p);
                continue;
            }
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                Entry& e = table_[p.label];
                e.ms += ms;
                e.calls++;
            }
            pool_.push_back(p.a);
            pool_.push_back(p.b);
        }
        pending_.swap(keep);
    }

    const std::map<std::string, Entry>& table() const { return table_; }
    void clear() { table_.clear(); }

    ~StageClock()
    {
        for (Pending& p : pending_) {
            (void) hipEventDestroy(p.a);
            (void) hipEventDestroy(p.b);
        }
        for (hipEvent_t e : pool_) (void) hipEventDestroy(e);
    }

private:
    struct Pending {
        std::string label;
        hipEvent_t a, b;
    };
    hipEvent_t take()
    {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void) hipEventCreate(&e);
        return e;
    }

    hipStream_t stream_ = nullptr;
    bool on_ = false;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, Entry> table_;
};

// One traced stage: a roctx range `range` and, when the clock is on, a timed stage `label`.
class Stage {
public:
    Stage(StageClock& c, const char* range, const char* label) : c_(c)
    {
        roctxRangePushA(range);
        token_ = label ? c_.start(label) : -1;
    }
    ~Stage()
    {
        c_.stop(token_);
        roctxRangePop();
    }
    Stage(const Stage&) = delete;
    Stage& operator=(const Stage&) = delete;

private:
    StageClock& c_;
    int token_;
};

}  // namespace armon
