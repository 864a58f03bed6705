// roctx ranges around the engine's phases (session set-up and runs, batch
// and family solves, certification, training; the facade's presolve, scaling,
// engine, postsolve and verification): a rocprofv3 --marker-trace timeline
// shows them beside the kernels. Without a profiler attached a range is a
// call into an idle library.
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>

namespace nir {

struct TraceRange {
    explicit TraceRange(const char* what) { roctxRangePushA(what); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace nir
