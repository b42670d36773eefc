// Drop-in for the reference's timer macros (Profiling.H:326-329): without
// ENABLE_TIMERS they expand to nothing there too.  GPU time is measured with
// HIP events (vt_step_timed) and rocprofv3 instead.
#ifndef PROFILING_H
#define PROFILING_H
#define PROFILE_DECLARE(name)
#define PROFILE_SCOPE(name)
#endif  // PROFILING_H
