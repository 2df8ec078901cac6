// Gather pass: per drawable handle resolution and indirect-command /
// DrawablePointers fill.
//
// Semantics follow src/CadR/shaders/processDrawables.comp:92-113 (main) and
// :77-89 (lookupHandle); the launch replaces the 32768 x Y split dispatch of
// src/CadR/Renderer.cpp:684-692 with one 1-D grid of wave64-friendly blocks
// (the split only works around Vulkan's per-dimension workgroup-count limit,
// the drawable -> index mapping is the same flat index).
//
// gfx950 shape: one lane per drawable, 256 lanes per block (4 waves), the
// handle-table depth a template parameter so each variant is a straight-line
// chain of dependent 8-byte loads.  The 48-byte record is read as three
// 16-byte loads, the 16-byte command and 32-byte pointer records are written
// with 16-byte stores; consecutive lanes touch consecutive records, so every
// wave reads and writes whole contiguous spans.
#include "internal.h"
#include "launch.h"
#include "handles.h"

namespace cadr {

template <int LEVEL>
__global__ __launch_bounds__(256) void processDrawablesKernel(const uint64_t* __restrict__ handleTable,
                                                              const uint4* __restrict__ drawableList,
                                                              uint4* __restrict__ indirectData,
                                                              uint4* __restrict__ drawablePointers,
                                                              uint64_t numDrawables)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i >= numDrawables) return;

	// DrawableGpuData as three 16-byte words
	const uint4 w0 = drawableList[3 * i + 0];  // vertexDataHandle, indexDataHandle
	const uint4 w1 = drawableList[3 * i + 1];  // matrixListHandle, drawableDataHandle
	const uint4 w2 = drawableList[3 * i + 2];  // primitiveSetHandle, primitiveSetOffset, padding
	const uint64_t vertexH = uint64_t(w0.x) | (uint64_t(w0.y) << 32);
	const uint64_t indexH = uint64_t(w0.z) | (uint64_t(w0.w) << 32);
	const uint64_t matrixListH = uint64_t(w1.x) | (uint64_t(w1.y) << 32);
	const uint64_t drawableDataH = uint64_t(w1.z) | (uint64_t(w1.w) << 32);
	const uint64_t primitiveSetH = uint64_t(w2.x) | (uint64_t(w2.y) << 32);
	const uint32_t primitiveSetOffset = w2.z;

	// PrimitiveSet {count, first} and MatrixList header
	const uint64_t psBase = lookupHandle<LEVEL>(handleTable, primitiveSetH);
	const uint64_t ml = lookupHandle<LEVEL>(handleTable, matrixListH);
	uint32_t psCount = 0, psFirst = 0, numMatrices = 0;
	if (psBase) {
		const GlobalConst<uint32_t> ps = globalAt<uint32_t>(psBase + primitiveSetOffset);
		psCount = ps[0];
		psFirst = ps[1];
	}
	if (ml) numMatrices = globalAt<uint32_t>(ml)[0];

	// VkDrawIndirectCommand {vertexCount, instanceCount, firstVertex, firstInstance}
	indirectData[i] = make_uint4(psCount, numMatrices, psFirst, 0u);

	// DrawablePointers {vertexDataPtr, indexDataPtr, matrixListPtr, drawableDataPtr}
	const uint64_t vertexPtr = lookupHandle<LEVEL>(handleTable, vertexH);
	const uint64_t indexPtr = lookupHandle<LEVEL>(handleTable, indexH);
	const uint64_t drawableDataPtr = lookupHandle<LEVEL>(handleTable, drawableDataH);
	drawablePointers[2 * i + 0] =
	    make_uint4(uint32_t(vertexPtr), uint32_t(vertexPtr >> 32), uint32_t(indexPtr), uint32_t(indexPtr >> 32));
	drawablePointers[2 * i + 1] = make_uint4(uint32_t(ml), uint32_t(ml >> 32), uint32_t(drawableDataPtr),
	                                         uint32_t(drawableDataPtr >> 32));
}

// Delta upload of the drawable list: list[slots[i]] = records[i], one lane
// per record, each 48-byte record moved as three 16-byte loads and stores.
__global__ __launch_bounds__(256) void scatterDrawablesKernel(const uint4* __restrict__ records,
                                                              const uint32_t* __restrict__ slots, uint32_t count,
                                                              uint4* __restrict__ list, uint64_t listSlots)
{
	const uint32_t i = blockIdx.x * 256u + threadIdx.x;
	if (i >= count) return;
	const uint64_t slot = slots[i];
	if (slot >= listSlots) return;
	const uint4 a = records[3ull * i + 0], b = records[3ull * i + 1], c = records[3ull * i + 2];
	list[3 * slot + 0] = a;
	list[3 * slot + 1] = b;
	list[3 * slot + 2] = c;
}

// Compacted DrawablePointers of a cull output: `split` workgroups per
// (view, StateSet) segment, workgroup c taking slots c*256 + t, stepping by
// split*256 (the counts are only known on the device, so every segment gets
// the same number of workgroups, enough for one of average size); slot j of
// the segment takes the 32-byte record of the drawable its command names
// (firstInstance, dword 4 of the 20-byte command), or zeros for an index
// past the drawable count.
__global__ __launch_bounds__(256) void compactPointersKernel(const uint32_t* __restrict__ commands,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ stateSetBase,
                                                             uint32_t numStateSets, uint32_t commandStride,
                                                             uint32_t split, const uint4* __restrict__ pointersIn,
                                                             uint64_t numDrawables, uint4* __restrict__ pointersOut)
{
	const uint32_t seg = blockIdx.x / split, c = blockIdx.x % split;
	const uint32_t w = seg / numStateSets, k = seg % numStateSets;
	const uint32_t base = stateSetBase[k];
	if (base >= commandStride) return;
	const uint32_t n = min(counts[uint64_t(w) * numStateSets + k], commandStride - base);
	const uint64_t first = uint64_t(w) * commandStride + base;
	for (uint32_t j = c * 256u + threadIdx.x; j < n; j += split * 256u) {
		const uint64_t slot = first + j;
		const uint32_t i = commands[5 * slot + 4];
		uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;
		if (i < numDrawables) {
			a = pointersIn[2ull * i + 0];
			b = pointersIn[2ull * i + 1];
		}
		pointersOut[2 * slot + 0] = a;
		pointersOut[2 * slot + 1] = b;
	}
}

}  // namespace cadr

using namespace cadr;

CADR_API int cadrCompactDrawablePointers(uint64_t commands, uint64_t counts, uint64_t stateSetBase, uint32_t numViews,
                                         uint32_t numStateSets, uint32_t commandStride, uint64_t pointersIn,
                                         uint64_t numDrawables, uint64_t pointersOut, cadrStream stream)
{
	if (numViews == 0 || numStateSets == 0) CADR_RETURN_NO_LAUNCH(stream);
	CADR_REQUIRE(commands && counts && stateSetBase && pointersIn && pointersOut,
	             "cadrCompactDrawablePointers: null device address");
	CADR_REQUIRE((commands % 4) == 0 && (counts % 4) == 0 && (stateSetBase % 4) == 0 && (pointersIn % 16) == 0 &&
	                 (pointersOut % 16) == 0,
	             "cadrCompactDrawablePointers: pointer arrays must be 16-byte aligned, the others 4-byte aligned");
	// workgroups per segment: one per 2048 slots of an average segment
	const uint64_t avg = (uint64_t(commandStride) + numStateSets - 1) / numStateSets;
	const uint32_t split = uint32_t(std::max<uint64_t>(1, (avg + 2047) / 2048));
	CADR_REQUIRE(uint64_t(numViews) * numStateSets * split < (1ull << 31),
	             "cadrCompactDrawablePointers: too many (view, StateSet) segments");
	launchKernel(compactPointersKernel, dim3(numViews * numStateSets * split), dim3(256), hipStream_t(stream),
	             reinterpret_cast<const uint32_t*>(commands), reinterpret_cast<const uint32_t*>(counts),
	             reinterpret_cast<const uint32_t*>(stateSetBase), numStateSets, commandStride, split,
	             reinterpret_cast<const uint4*>(pointersIn), numDrawables, reinterpret_cast<uint4*>(pointersOut));
	CADR_CHECK_HIP(hipGetLastError(), "compactPointersKernel launch");
	return ok();
}

CADR_API int cadrScatterDrawables(uint64_t drawableList, uint64_t listSlots, uint64_t records, uint64_t slots,
                                  uint32_t count, cadrStream stream)
{
	if (count == 0) return ok();
	CADR_REQUIRE(drawableList && records && slots, "cadrScatterDrawables: null device address");
	CADR_REQUIRE((drawableList % 16) == 0 && (records % 16) == 0 && (slots % 4) == 0,
	             "cadrScatterDrawables: drawableList and records must be 16-byte aligned, slots 4-byte aligned");
	hipLaunchKernelGGL(scatterDrawablesKernel, dim3((count + 255u) / 256u), dim3(256), 0, hipStream_t(stream),
	                   reinterpret_cast<const uint4*>(records), reinterpret_cast<const uint32_t*>(slots), count,
	                   reinterpret_cast<uint4*>(drawableList), listSlots);
	CADR_CHECK_HIP(hipGetLastError(), "scatterDrawablesKernel launch");
	return ok();
}

CADR_API int cadrProcessDrawables(const cadrProcessDrawablesPushConstants* pc, uint32_t handleLevel,
                                  uint64_t numDrawables, cadrStream stream)
{
	CADR_REQUIRE(pc, "cadrProcessDrawables: pushConstants is NULL");
	if (numDrawables == 0) CADR_RETURN_NO_LAUNCH(stream);  // Renderer.cpp:600-620 returns early
	CADR_REQUIRE(numDrawables <= CADR_MAX_DRAWABLES,
	             "cadrProcessDrawables: numDrawables exceeds the 1Gi limit (Renderer.cpp:687)");
	CADR_REQUIRE(handleLevel >= 1 && handleLevel <= 3, "cadrProcessDrawables: handleLevel must be 1, 2 or 3");
	CADR_REQUIRE(pc->handleTable && pc->drawableList && pc->indirectData && pc->drawablePointers,
	             "cadrProcessDrawables: null device address in push constants");
	CADR_REQUIRE((pc->handleTable % 8) == 0, "cadrProcessDrawables: handleTable must be 8-byte aligned");
	CADR_REQUIRE((pc->drawableList % 16) == 0 && (pc->indirectData % 16) == 0 && (pc->drawablePointers % 16) == 0,
	             "cadrProcessDrawables: drawableList, indirectData and drawablePointers must be 16-byte aligned");

	const uint64_t blocks = (numDrawables + 255) / 256;
	const auto* table = reinterpret_cast<const uint64_t*>(pc->handleTable);
	const auto* list = reinterpret_cast<const uint4*>(pc->drawableList);
	auto* indirect = reinterpret_cast<uint4*>(pc->indirectData);
	auto* pointers = reinterpret_cast<uint4*>(pc->drawablePointers);
	hipStream_t s = hipStream_t(stream);
	auto kernel = handleLevel == 1 ? processDrawablesKernel<1>
	              : handleLevel == 2 ? processDrawablesKernel<2> : processDrawablesKernel<3>;
	launchKernel(kernel, dim3(uint32_t(blocks)), dim3(256), s, table, list, indirect, pointers, numDrawables);
	CADR_CHECK_HIP(hipGetLastError(), "processDrawablesKernel launch");
	return ok();
}
