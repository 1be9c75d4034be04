// Optional per-kernel-class timing with HIP events on the launch stream.
//
// Off by default (one branch per launch).  moao_hip_profile_begin() clears and enables it;
// every instrumented launch then records an event before and after itself on the stream it is
// launched on and its algorithmic work; moao_hip_profile_end() synchronises and sums the
// durations per class.  Events come from a pool that is reused across sessions.
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

enum ProfClass { kProfGemmF64 = 0, kProfGemmF32, kProfLeaf, kProfMatcov, kProfConvert,
                 kProfFft, kProfOther, kProfClasses };

bool profiling();
// Returns an index to pass to prof_stop, or -1 when profiling is off.
int prof_start(int cls, double work, hipStream_t s);
void prof_stop(int idx, hipStream_t s);

struct ProfScope {
    int idx;
    hipStream_t s;
    ProfScope(int cls, double work, hipStream_t st) : idx(prof_start(cls, work, st)), s(st) {}
    ~ProfScope() { prof_stop(idx, s); }
};

}  // namespace moao
