// Internal helpers shared by the translation units of libcadr_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cadr_hip.h"

namespace cadr {

// Records the detail text of the calling thread's last error and returns code.
int setError(int code, const std::string& message);
int setHipError(hipError_t err, const char* what);
inline int ok() { return CADR_SUCCESS; }

// Compute-unit count of the current device (cached per device).
int computeUnitCount();

// Whether the cull entry point of `path` has a kernel for this shape.
enum class CullPath { View, Batch, Fused, FusedBatch };
bool cullHasKernel(const cadrCullConfig* c, CullPath path);

}  // namespace cadr

#define CADR_CHECK_HIP(expr, what)                                  \
	do {                                                            \
		hipError_t e_ = (expr);                                     \
		if (e_ != hipSuccess) return ::cadr::setHipError(e_, what); \
	} while (0)

#define CADR_REQUIRE(cond, msg)                                                     \
	do {                                                                            \
		if (!(cond)) return ::cadr::setError(CADR_ERROR_INVALID_VALUE, msg);       \
	} while (0)

#define CADR_API extern "C" __attribute__((visibility("default")))
