// Kernel launches that can carry the calling thread's timing events
// (cadrTimeNextLaunch).  Included by the HIP translation units only.
#pragma once

#include <hip/hip_ext.h>

#include "internal.h"

namespace cadr {

// The start / stop events a caller asked the next timed launch to record;
// taking them clears them.
struct LaunchEvents {
	hipEvent_t start = nullptr, stop = nullptr;
};
LaunchEvents takeLaunchEvents();

// Launches kernel<<<grid, block, 0, s>>>(args...).  When the calling thread
// has timing events pending, the launch records them itself: the events'
// timestamps are the dispatch's own start and end, with no marker commands
// queued around the kernel.
template <class F, class... Args>
inline void launchKernel(F kernel, dim3 grid, dim3 block, hipStream_t s, Args... args)
{
	const LaunchEvents e = takeLaunchEvents();
	if (e.start || e.stop)
		hipExtLaunchKernelGGL(kernel, grid, block, 0u, s, e.start, e.stop, 0u, args...);
	else
		hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// For an entry point that launches nothing this time: records pending timing
// events back to back on s, so that they measure zero instead of staying
// unrecorded.
hipError_t recordSkippedLaunch(hipStream_t s);

}  // namespace cadr

// Return path of an entry point that has nothing to launch this time.
#define CADR_RETURN_NO_LAUNCH(stream)                                                         \
	do {                                                                                  \
		CADR_CHECK_HIP(::cadr::recordSkippedLaunch(hipStream_t(stream)), "hipEventRecord"); \
		return ::cadr::ok();                                                              \
	} while (0)
