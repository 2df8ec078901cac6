// Device, memory, stream and event entry points of the C ABI, and the error
// bookkeeping shared by the other translation units.
//
// Stand-in for the Vulkan device plumbing of the reference
// (src/CadR/Renderer.cpp:97-313 init, :946-999 executeCopyOperations,
// src/CadR/DataMemory.cpp:255-397 block allocation).
#include "internal.h"
#include "launch.h"

#include <mutex>
#include <vector>

namespace cadr {

static thread_local std::string g_lastError;
static thread_local LaunchEvents g_launchEvents;

LaunchEvents takeLaunchEvents()
{
	const LaunchEvents e = g_launchEvents;
	g_launchEvents = LaunchEvents{};
	return e;
}

hipError_t recordSkippedLaunch(hipStream_t s)
{
	const LaunchEvents e = takeLaunchEvents();
	hipError_t err = hipSuccess;
	if (e.start) err = hipEventRecord(e.start, s);
	if (err == hipSuccess && e.stop) err = hipEventRecord(e.stop, s);
	return err;
}

int setError(int code, const std::string& message)
{
	g_lastError = message;
	return code;
}

int setHipError(hipError_t err, const char* what)
{
	int code = (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
	               ? CADR_ERROR_OUT_OF_RESOURCES
	               : (err == hipErrorNotReady ? CADR_ERROR_NOT_READY : CADR_ERROR_DEVICE);
	return setError(code, std::string(what) + ": " + hipGetErrorString(err));
}

int computeUnitCount()
{
	static std::mutex m;
	static std::vector<int> cache;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 0;
	std::lock_guard<std::mutex> lock(m);
	if (int(cache.size()) <= dev) cache.resize(dev + 1, 0);
	if (cache[dev] == 0) {
		hipDeviceProp_t p;
		if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
		cache[dev] = p.multiProcessorCount;
	}
	return cache[dev];
}

// Grid-stride fill of 32-bit words; used instead of hipMemsetAsync so that
// the clear can be recorded into a HIP graph.
__global__ void fillU32Kernel(uint32_t* __restrict__ dst, uint32_t value, uint64_t n)
{
	uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		dst[i] = value;
}

}  // namespace cadr

using namespace cadr;

CADR_API const char* cadrGetResultString(int result)
{
	switch (result) {
	case CADR_SUCCESS: return "CADR_SUCCESS";
	case CADR_ERROR_INVALID_VALUE: return "CADR_ERROR_INVALID_VALUE";
	case CADR_ERROR_OUT_OF_RESOURCES: return "CADR_ERROR_OUT_OF_RESOURCES";
	case CADR_ERROR_TIMEOUT: return "CADR_ERROR_TIMEOUT";
	case CADR_ERROR_DEVICE: return "CADR_ERROR_DEVICE";
	case CADR_ERROR_NOT_READY: return "CADR_ERROR_NOT_READY";
	default: return "CADR_ERROR_UNKNOWN";
	}
}

CADR_API const char* cadrGetLastErrorMessage(void) { return g_lastError.c_str(); }

CADR_API int cadrSetDevice(int deviceIndex)
{
	CADR_CHECK_HIP(hipSetDevice(deviceIndex), "hipSetDevice");
	return ok();
}

CADR_API int cadrGetDeviceCount(int* count)
{
	CADR_REQUIRE(count, "cadrGetDeviceCount: count is NULL");
	CADR_CHECK_HIP(hipGetDeviceCount(count), "hipGetDeviceCount");
	return ok();
}

CADR_API int cadrGetComputeUnitCount(int* count)
{
	CADR_REQUIRE(count, "cadrGetComputeUnitCount: count is NULL");
	*count = computeUnitCount();
	if (*count <= 0) return setError(CADR_ERROR_DEVICE, "cannot query the compute unit count");
	return ok();
}

CADR_API int cadrDeviceMalloc(uint64_t numBytes, uint64_t* deviceAddress)
{
	CADR_REQUIRE(deviceAddress, "cadrDeviceMalloc: deviceAddress is NULL");
	*deviceAddress = 0;
	if (numBytes == 0) return ok();
	void* p = nullptr;
	hipError_t e = hipMalloc(&p, numBytes);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		return setError(CADR_ERROR_OUT_OF_RESOURCES,
		                "cadrDeviceMalloc: cannot allocate " + std::to_string(numBytes) + " bytes");
	}
	*deviceAddress = reinterpret_cast<uint64_t>(p);
	return ok();
}

CADR_API int cadrDeviceFree(uint64_t deviceAddress)
{
	if (deviceAddress == 0) return ok();
	CADR_CHECK_HIP(hipFree(reinterpret_cast<void*>(deviceAddress)), "hipFree");
	return ok();
}

CADR_API int cadrHostMallocPinned(uint64_t numBytes, void** hostPointer)
{
	CADR_REQUIRE(hostPointer, "cadrHostMallocPinned: hostPointer is NULL");
	*hostPointer = nullptr;
	if (numBytes == 0) return ok();
	hipError_t e = hipHostMalloc(hostPointer, numBytes, hipHostMallocDefault);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		return setError(CADR_ERROR_OUT_OF_RESOURCES, "cadrHostMallocPinned: cannot allocate pinned memory");
	}
	return ok();
}

CADR_API int cadrHostFreePinned(void* hostPointer)
{
	if (!hostPointer) return ok();
	CADR_CHECK_HIP(hipHostFree(hostPointer), "hipHostFree");
	return ok();
}

CADR_API int cadrMemcpyHtoD(uint64_t dst, const void* src, uint64_t numBytes, cadrStream stream)
{
	if (numBytes == 0) return ok();
	CADR_REQUIRE(dst && src, "cadrMemcpyHtoD: null pointer");
	CADR_CHECK_HIP(hipMemcpyAsync(reinterpret_cast<void*>(dst), src, numBytes, hipMemcpyHostToDevice,
	                              hipStream_t(stream)),
	               "hipMemcpyAsync(HtoD)");
	return ok();
}

CADR_API int cadrMemcpyDtoH(void* dst, uint64_t src, uint64_t numBytes, cadrStream stream)
{
	if (numBytes == 0) return ok();
	CADR_REQUIRE(dst && src, "cadrMemcpyDtoH: null pointer");
	CADR_CHECK_HIP(hipMemcpyAsync(dst, reinterpret_cast<const void*>(src), numBytes, hipMemcpyDeviceToHost,
	                              hipStream_t(stream)),
	               "hipMemcpyAsync(DtoH)");
	return ok();
}

CADR_API int cadrMemcpyDtoD(uint64_t dst, uint64_t src, uint64_t numBytes, cadrStream stream)
{
	if (numBytes == 0) return ok();
	CADR_REQUIRE(dst && src, "cadrMemcpyDtoD: null pointer");
	CADR_CHECK_HIP(hipMemcpyAsync(reinterpret_cast<void*>(dst), reinterpret_cast<const void*>(src), numBytes,
	                              hipMemcpyDeviceToDevice, hipStream_t(stream)),
	               "hipMemcpyAsync(DtoD)");
	return ok();
}

CADR_API int cadrFillU32(uint64_t dst, uint32_t value, uint64_t numBytes, cadrStream stream)
{
	if (numBytes == 0) return ok();
	CADR_REQUIRE(dst && (dst % 4) == 0 && (numBytes % 4) == 0, "cadrFillU32: unaligned address or size");
	uint64_t n = numBytes / 4;
	int cu = computeUnitCount();
	uint64_t blocks = (n + 255) / 256;
	uint64_t maxBlocks = uint64_t(cu > 0 ? cu : 256) * 8;
	if (blocks > maxBlocks) blocks = maxBlocks;
	hipLaunchKernelGGL(fillU32Kernel, dim3(uint32_t(blocks)), dim3(256), 0, hipStream_t(stream),
	                   reinterpret_cast<uint32_t*>(dst), value, n);
	CADR_CHECK_HIP(hipGetLastError(), "fillU32Kernel launch");
	return ok();
}

CADR_API int cadrStreamCreate(cadrStream* stream)
{
	CADR_REQUIRE(stream, "cadrStreamCreate: stream is NULL");
	hipStream_t s = nullptr;
	CADR_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
	*stream = s;
	return ok();
}

CADR_API int cadrStreamDestroy(cadrStream stream)
{
	if (!stream) return ok();
	CADR_CHECK_HIP(hipStreamDestroy(hipStream_t(stream)), "hipStreamDestroy");
	return ok();
}

CADR_API int cadrStreamSynchronize(cadrStream stream)
{
	CADR_CHECK_HIP(hipStreamSynchronize(hipStream_t(stream)), "hipStreamSynchronize");
	return ok();
}

CADR_API int cadrDeviceSynchronize(void)
{
	CADR_CHECK_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
	return ok();
}

CADR_API int cadrEventCreate(cadrEvent* event)
{
	CADR_REQUIRE(event, "cadrEventCreate: event is NULL");
	hipEvent_t e = nullptr;
	CADR_CHECK_HIP(hipEventCreate(&e), "hipEventCreate");
	*event = e;
	return ok();
}

CADR_API int cadrEventDestroy(cadrEvent event)
{
	if (!event) return ok();
	CADR_CHECK_HIP(hipEventDestroy(hipEvent_t(event)), "hipEventDestroy");
	return ok();
}

CADR_API int cadrEventRecord(cadrEvent event, cadrStream stream)
{
	CADR_REQUIRE(event, "cadrEventRecord: event is NULL");
	CADR_CHECK_HIP(hipEventRecord(hipEvent_t(event), hipStream_t(stream)), "hipEventRecord");
	return ok();
}

CADR_API int cadrEventSynchronize(cadrEvent event)
{
	CADR_REQUIRE(event, "cadrEventSynchronize: event is NULL");
	CADR_CHECK_HIP(hipEventSynchronize(hipEvent_t(event)), "hipEventSynchronize");
	return ok();
}

CADR_API int cadrEventQuery(cadrEvent event)
{
	CADR_REQUIRE(event, "cadrEventQuery: event is NULL");
	hipError_t e = hipEventQuery(hipEvent_t(event));
	if (e == hipErrorNotReady) return CADR_ERROR_NOT_READY;
	CADR_CHECK_HIP(e, "hipEventQuery");
	return ok();
}

CADR_API int cadrTimeNextLaunch(cadrEvent start, cadrEvent stop)
{
	g_launchEvents = LaunchEvents{hipEvent_t(start), hipEvent_t(stop)};
	return ok();
}

CADR_API int cadrEventElapsedTime(float* milliseconds, cadrEvent start, cadrEvent stop)
{
	CADR_REQUIRE(milliseconds && start && stop, "cadrEventElapsedTime: null argument");
	CADR_CHECK_HIP(hipEventElapsedTime(milliseconds, hipEvent_t(start), hipEvent_t(stop)),
	               "hipEventElapsedTime");
	return ok();
}
