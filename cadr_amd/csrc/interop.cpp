// Device buffers shared with a Vulkan renderer through dma-buf file
// descriptors (VK_KHR_external_memory_fd), so that the renderer's
// vkCmdDrawIndexedIndirectCount and its shaders' DrawablePointers[gl_DrawID]
// read the cull's outputs where the cull writes them.
//
// The reference allocates its indirect and pointer buffers as Vulkan buffers
// with device addresses (src/CadR/Renderer.cpp:446-595) and its shaders read
// them by gl_DrawID (examples/RenderingPerformance/shader.vert:99-123).  Here
// either side may own the memory:
//   cadrSharedBufferCreate - allocated here (HIP virtual memory, a physical
//     allocation that can be exported), mapped at a device address the cull
//     writes to; cadrSharedBufferExportFd hands the renderer a descriptor for
//     vkAllocateMemory with VkImportMemoryFdInfoKHR;
//   cadrSharedBufferImport - memory the renderer allocated and exported
//     (vkGetMemoryFdKHR), mapped here at a device address.
#include <unistd.h>

#include <algorithm>

#include "internal.h"

using namespace cadr;

struct cadrSharedBuffer {
	int device = 0;
	uint64_t size = 0;
	void* address = nullptr;
	bool imported = false;
	hipMemGenericAllocationHandle_t handle{};  // created buffers
	hipExternalMemory_t external = nullptr;    // imported buffers
};

namespace {

hipMemAllocationProp exportableProp(int device)
{
	hipMemAllocationProp prop{};
	prop.type = hipMemAllocationTypePinned;
	prop.location.type = hipMemLocationTypeDevice;
	prop.location.id = device;
	prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
	return prop;
}

}  // namespace

CADR_API int cadrSharedBufferCreate(uint64_t size, cadrSharedBuffer** buffer)
{
	CADR_REQUIRE(buffer, "cadrSharedBufferCreate: buffer is NULL");
	*buffer = nullptr;
	CADR_REQUIRE(size > 0, "cadrSharedBufferCreate: size is 0");
	auto* b = new cadrSharedBuffer;
	auto fail = [&](hipError_t e, const char* what) {
		cadrSharedBufferDestroy(b);
		return setHipError(e, what);
	};
	hipError_t e = hipGetDevice(&b->device);
	if (e != hipSuccess) return fail(e, "hipGetDevice");
	const hipMemAllocationProp prop = exportableProp(b->device);
	size_t granularity = 0;
	e = hipMemGetAllocationGranularity(&granularity, &prop, hipMemAllocationGranularityMinimum);
	if (e != hipSuccess) return fail(e, "hipMemGetAllocationGranularity");
	granularity = std::max<size_t>(granularity, 1);
	const uint64_t rounded = (size + granularity - 1) / granularity * granularity;
	e = hipMemCreate(&b->handle, rounded, &prop, 0);
	if (e != hipSuccess) return fail(e, "hipMemCreate");
	b->size = rounded;  // from here on Destroy releases the handle
	e = hipMemAddressReserve(&b->address, rounded, 0, nullptr, 0);
	if (e != hipSuccess) {
		b->address = nullptr;
		return fail(e, "hipMemAddressReserve");
	}
	e = hipMemMap(b->address, rounded, 0, b->handle, 0);
	if (e != hipSuccess) return fail(e, "hipMemMap");
	hipMemAccessDesc access{};
	access.location = prop.location;
	access.flags = hipMemAccessFlagsProtReadWrite;
	e = hipMemSetAccess(b->address, rounded, &access, 1);
	if (e != hipSuccess) return fail(e, "hipMemSetAccess");
	*buffer = b;
	return ok();
}

CADR_API int cadrSharedBufferImport(int fd, uint64_t size, cadrSharedBuffer** buffer)
{
	CADR_REQUIRE(buffer, "cadrSharedBufferImport: buffer is NULL");
	*buffer = nullptr;
	CADR_REQUIRE(fd >= 0 && size > 0, "cadrSharedBufferImport: invalid descriptor or size");
	auto* b = new cadrSharedBuffer;
	b->imported = true;
	auto fail = [&](hipError_t e, const char* what) {
		cadrSharedBufferDestroy(b);
		return setHipError(e, what);
	};
	hipError_t e = hipGetDevice(&b->device);
	if (e != hipSuccess) return fail(e, "hipGetDevice");
	hipExternalMemoryHandleDesc desc{};
	desc.type = hipExternalMemoryHandleTypeOpaqueFd;
	desc.handle.fd = fd;
	desc.size = size;
	e = hipImportExternalMemory(&b->external, &desc);
	if (e != hipSuccess) {
		b->external = nullptr;
		return fail(e, "hipImportExternalMemory");
	}
	hipExternalMemoryBufferDesc range{};
	range.offset = 0;
	range.size = size;
	e = hipExternalMemoryGetMappedBuffer(&b->address, b->external, &range);
	if (e != hipSuccess) {
		b->address = nullptr;
		return fail(e, "hipExternalMemoryGetMappedBuffer");
	}
	b->size = size;
	*buffer = b;
	return ok();
}

CADR_API uint64_t cadrSharedBufferAddress(const cadrSharedBuffer* buffer)
{
	return buffer ? reinterpret_cast<uint64_t>(buffer->address) : 0u;
}

CADR_API uint64_t cadrSharedBufferSize(const cadrSharedBuffer* buffer) { return buffer ? buffer->size : 0u; }

CADR_API int cadrSharedBufferExportFd(const cadrSharedBuffer* buffer, int* fd)
{
	CADR_REQUIRE(buffer && fd, "cadrSharedBufferExportFd: null argument");
	CADR_REQUIRE(!buffer->imported, "cadrSharedBufferExportFd: an imported buffer is exported by its owner");
	*fd = -1;
	int out = -1;
	CADR_CHECK_HIP(hipMemExportToShareableHandle(&out, buffer->handle, hipMemHandleTypePosixFileDescriptor, 0),
	               "hipMemExportToShareableHandle");
	*fd = out;
	return ok();
}

CADR_API void cadrSharedBufferDestroy(cadrSharedBuffer* buffer)
{
	if (!buffer) return;
	if (buffer->imported) {
		// the whole-range mapping goes with the external-memory object
		if (buffer->external) (void)hipDestroyExternalMemory(buffer->external);
	} else if (buffer->size) {
		if (buffer->address) {
			(void)hipMemUnmap(buffer->address, buffer->size);
			(void)hipMemAddressFree(buffer->address, buffer->size);
		}
		(void)hipMemRelease(buffer->handle);
	}
	(void)hipGetLastError();
	delete buffer;
}
