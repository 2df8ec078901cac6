// Cull inputs from the reference's own records, and the reference's draw
// contract for the compacted commands: the device side of the CadR facade's
// cull frame (Renderer::recordDrawableCulling).
//
// The reference binds, per drawable, a Geometry's PrimitiveSet allocation at a
// byte offset and a MatrixList (DrawableGpuData, src/CadR/Drawable.h:32-43) and
// its consumer reads DrawablePointers[gl_DrawID] and
// matrices[gl_InstanceIndex] (examples/RenderingPerformance/shader.vert:
// 99-123).  Three kernels connect that layout to the cull pass:
//   buildMeshTableKernel - one cadrMesh per (Geometry, PrimitiveSet offset):
//     LOD k is the PrimitiveSet record at lookup(handle) + offset + 8k, read
//     as processDrawables.comp:97 reads LOD 0;
//   cullItemsKernel - one lane per cull item (one matrix of a drawable's
//     MatrixList): resolves the MatrixList handle and writes the item's world
//     sphere (src/CadR/BoundingSphere.h:70-87) and the mesh index the facade
//     keeps in the record's padding word;
//   finishCullKernel - per compacted command: firstInstance (the cull's item
//     index) becomes the matrix index inside the drawable's MatrixList, and
//     the drawable's DrawablePointers record is copied to the command's slot.
#include "internal.h"
#include "fp32.h"
#include "handles.h"
#include "matrix_io.h"

#include <algorithm>

namespace cadr {

static_assert(sizeof(cadrMeshSource) == 32, "cadrMeshSource is 32 bytes");
static_assert(sizeof(cadrMesh) == 64, "cadrMesh is 64 bytes");

template <int LEVEL>
__global__ __launch_bounds__(256) void buildMeshTableKernel(const uint64_t* __restrict__ root,
                                                            const cadrMeshSource* __restrict__ sources,
                                                            uint32_t count, cadrMesh* __restrict__ meshes)
{
	const uint32_t m = blockIdx.x * 256u + threadIdx.x;
	if (m >= count) return;
	const cadrMeshSource s = sources[m];
	cadrMesh o{};
	const uint64_t ps = s.lodCount ? lookupHandle<LEVEL>(root, s.primitiveSetHandle) : 0;
	if (ps == 0) {
		o.radius = -__builtin_inff();  // BoundingSphere::empty(): never visible
		o.lodCount = 1;
	} else {
		o.center[0] = s.sphere[0];
		o.center[1] = s.sphere[1];
		o.center[2] = s.sphere[2];
		o.radius = s.sphere[3];
		const uint32_t lods = s.lodCount <= 4u ? s.lodCount : 4u;
		o.lodCount = lods;
		for (uint32_t k = 0; k < lods; k++) {
			const GlobalConst<uint32_t> rec = globalAt<uint32_t>(ps + s.primitiveSetOffset + 8ull * k);
			o.indexCount[k] = rec[0];
			o.firstIndex[k] = rec[1];
		}
	}
	meshes[m] = o;
}

struct CullItemsArgs {
	const uint64_t* root;
	const uint4* list;
	const uint32_t* itemBase;
	const uint32_t* itemDrawable;
	const cadrMesh* meshes;
	float4* spheres;
	uint32_t* meshIndex;
	uint64_t numDrawables;
	uint64_t numItems;
	uint32_t numMeshes;
};

// One lane per item: consecutive items of one drawable read consecutive
// matrices of its MatrixList, and neighbouring lanes share the drawable's
// record, handle-table walk and list header in the caches.
template <int LEVEL>
__global__ __launch_bounds__(256) void cullItemsKernel(CullItemsArgs a)
{
	const uint64_t it = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (it >= a.numItems) return;
	uint64_t i = it;
	uint32_t j = 0;
	if (a.itemBase) {
		i = globalAt(a.itemDrawable)[it];
		j = i < a.numDrawables ? uint32_t(it - globalAt(a.itemBase)[i]) : 0u;
	}
	float4 s = make_float4(0.f, 0.f, 0.f, -__builtin_inff());
	uint32_t mesh = ~0u;
	if (i < a.numDrawables) {
		const GlobalConst<uint4> list = globalAt(a.list);
		const uint4 w1 = loadU4(list + (3 * i + 1));  // matrixListHandle, drawableDataHandle
		const uint4 w2 = loadU4(list + (3 * i + 2));  // primitiveSetHandle, primitiveSetOffset, padding
		const uint64_t ml = lookupHandle<LEVEL>(a.root, uint64_t(w1.x) | (uint64_t(w1.y) << 32));
		const uint32_t numMatrices = ml ? globalAt<uint32_t>(ml)[0] : 0u;
		mesh = w2.w;
		if (j < numMatrices && mesh < a.numMeshes) {
			// the mesh's first 16 bytes: local centre and radius
			const float4 local = loadF4(globalAt<float4>(reinterpret_cast<uint64_t>(a.meshes + mesh)));
			const GlobalConst<float4> matrices = globalAt<float4>(ml + 64);
			float4 M[4];
#pragma unroll
			for (int c = 0; c < 4; c++) M[c] = loadF4(matrices + (4ull * j + c));
			s = sphereTransform(M, local);
		}
	}
	storeF4(globalMutAt(a.spheres) + it, s);
	globalMutAt(a.meshIndex)[it] = mesh;
}

// `split` workgroups per (view, StateSet) segment, as compactPointersKernel.
__global__ __launch_bounds__(256) void finishCullKernel(uint32_t* __restrict__ commands,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ stateSetBase,
                                                        uint32_t numStateSets, uint32_t commandStride, uint32_t split,
                                                        const uint32_t* __restrict__ itemBase,
                                                        const uint32_t* __restrict__ itemDrawable, uint64_t numItems,
                                                        const uint4* __restrict__ pointersIn, uint64_t numDrawables,
                                                        uint4* __restrict__ pointersOut)
{
	const uint32_t seg = blockIdx.x / split, c = blockIdx.x % split;
	const uint32_t w = seg / numStateSets, k = seg % numStateSets;
	const uint32_t base = stateSetBase[k];
	if (base >= commandStride) return;
	const uint32_t n = min(counts[uint64_t(w) * numStateSets + k], commandStride - base);
	const uint64_t first = uint64_t(w) * commandStride + base;
	for (uint32_t j = c * 256u + threadIdx.x; j < n; j += split * 256u) {
		const uint64_t slot = first + j;
		const uint32_t item = commands[5 * slot + 4];
		uint64_t drawable = ~0ull;
		if (item < numItems) {
			uint32_t matrix = 0;
			drawable = item;
			if (itemBase) {
				drawable = itemDrawable[item];
				matrix = drawable < numDrawables ? item - itemBase[drawable] : 0u;
			}
			commands[5 * slot + 4] = matrix;
		}
		if (pointersOut) {
			uint4 p0 = make_uint4(0u, 0u, 0u, 0u), p1 = p0;
			if (drawable < numDrawables) {
				p0 = pointersIn[2ull * drawable + 0];
				p1 = pointersIn[2ull * drawable + 1];
			}
			pointersOut[2 * slot + 0] = p0;
			pointersOut[2 * slot + 1] = p1;
		}
	}
}

}  // namespace cadr

using namespace cadr;

CADR_API int cadrBuildMeshTable(uint64_t handleTable, uint32_t handleLevel, uint64_t sources, uint32_t count,
                                uint64_t meshes, cadrStream stream)
{
	if (count == 0) return ok();
	CADR_REQUIRE(handleLevel >= 1 && handleLevel <= 3, "cadrBuildMeshTable: handleLevel must be 1, 2 or 3");
	CADR_REQUIRE(handleTable && sources && meshes, "cadrBuildMeshTable: null device address");
	CADR_REQUIRE(handleTable % 8 == 0 && sources % 16 == 0 && meshes % 16 == 0,
	             "cadrBuildMeshTable: sources and meshes must be 16-byte aligned, handleTable 8-byte aligned");
	const auto* root = reinterpret_cast<const uint64_t*>(handleTable);
	const auto* src = reinterpret_cast<const cadrMeshSource*>(sources);
	auto* out = reinterpret_cast<cadrMesh*>(meshes);
	const dim3 grid((count + 255u) / 256u);
	hipStream_t s = hipStream_t(stream);
	switch (handleLevel) {
	case 1: hipLaunchKernelGGL(buildMeshTableKernel<1>, grid, dim3(256), 0, s, root, src, count, out); break;
	case 2: hipLaunchKernelGGL(buildMeshTableKernel<2>, grid, dim3(256), 0, s, root, src, count, out); break;
	default: hipLaunchKernelGGL(buildMeshTableKernel<3>, grid, dim3(256), 0, s, root, src, count, out); break;
	}
	CADR_CHECK_HIP(hipGetLastError(), "buildMeshTableKernel launch");
	return ok();
}

CADR_API int cadrBuildCullItems(const cadrCullItemsDesc* d, cadrStream stream)
{
	CADR_REQUIRE(d, "cadrBuildCullItems: desc is NULL");
	if (d->numItems == 0) return ok();
	CADR_REQUIRE(d->handleLevel >= 1 && d->handleLevel <= 3, "cadrBuildCullItems: handleLevel must be 1, 2 or 3");
	CADR_REQUIRE(d->handleTable && d->drawableList && d->spheres && d->meshIndex && (d->meshes || d->numMeshes == 0),
	             "cadrBuildCullItems: null device address");
	CADR_REQUIRE(d->handleTable % 8 == 0 && d->drawableList % 16 == 0 && d->spheres % 16 == 0 && d->meshes % 16 == 0 &&
	                 d->meshIndex % 4 == 0 && d->itemBase % 4 == 0 && d->itemDrawable % 4 == 0,
	             "cadrBuildCullItems: misaligned array");
	CADR_REQUIRE(d->itemBase ? d->itemDrawable != 0 : d->numItems == d->numDrawables,
	             "cadrBuildCullItems: itemBase needs itemDrawable; without them there is one item per drawable");
	CADR_REQUIRE(d->numDrawables < (1ull << 32) && d->numItems < (1ull << 32),
	             "cadrBuildCullItems: more than 2^32 - 1 drawables or items");
	CullItemsArgs a;
	a.root = reinterpret_cast<const uint64_t*>(d->handleTable);
	a.list = reinterpret_cast<const uint4*>(d->drawableList);
	a.itemBase = reinterpret_cast<const uint32_t*>(d->itemBase);
	a.itemDrawable = reinterpret_cast<const uint32_t*>(d->itemDrawable);
	a.meshes = reinterpret_cast<const cadrMesh*>(d->meshes);
	a.spheres = reinterpret_cast<float4*>(d->spheres);
	a.meshIndex = reinterpret_cast<uint32_t*>(d->meshIndex);
	a.numDrawables = d->numDrawables;
	a.numItems = d->numItems;
	a.numMeshes = d->numMeshes;
	const dim3 grid(uint32_t((d->numItems + 255u) / 256u));
	hipStream_t s = hipStream_t(stream);
	switch (d->handleLevel) {
	case 1: hipLaunchKernelGGL(cullItemsKernel<1>, grid, dim3(256), 0, s, a); break;
	case 2: hipLaunchKernelGGL(cullItemsKernel<2>, grid, dim3(256), 0, s, a); break;
	default: hipLaunchKernelGGL(cullItemsKernel<3>, grid, dim3(256), 0, s, a); break;
	}
	CADR_CHECK_HIP(hipGetLastError(), "cullItemsKernel launch");
	return ok();
}

CADR_API int cadrFinishCullCommands(uint64_t commands, uint64_t counts, uint64_t stateSetBase, uint32_t numViews,
                                    uint32_t numStateSets, uint32_t commandStride, uint64_t itemBase,
                                    uint64_t itemDrawable, uint64_t numItems, uint64_t pointersIn,
                                    uint64_t numDrawables, uint64_t pointersOut, cadrStream stream)
{
	if (numViews == 0 || numStateSets == 0) return ok();
	CADR_REQUIRE(commands && counts && stateSetBase && (!pointersOut || pointersIn) && (!itemBase || itemDrawable),
	             "cadrFinishCullCommands: null device address");
	CADR_REQUIRE(commands % 4 == 0 && counts % 4 == 0 && stateSetBase % 4 == 0 && itemBase % 4 == 0 &&
	                 itemDrawable % 4 == 0 && pointersIn % 16 == 0 && pointersOut % 16 == 0,
	             "cadrFinishCullCommands: pointer arrays must be 16-byte aligned, the others 4-byte aligned");
	const uint64_t avg = (uint64_t(commandStride) + numStateSets - 1) / numStateSets;
	const uint32_t split = uint32_t(std::max<uint64_t>(1, (avg + 2047) / 2048));
	CADR_REQUIRE(uint64_t(numViews) * numStateSets * split < (1ull << 31),
	             "cadrFinishCullCommands: too many (view, StateSet) segments");
	hipLaunchKernelGGL(finishCullKernel, dim3(numViews * numStateSets * split), dim3(256), 0, hipStream_t(stream),
	                   reinterpret_cast<uint32_t*>(commands), reinterpret_cast<const uint32_t*>(counts),
	                   reinterpret_cast<const uint32_t*>(stateSetBase), numStateSets, commandStride, split,
	                   reinterpret_cast<const uint32_t*>(itemBase), reinterpret_cast<const uint32_t*>(itemDrawable),
	                   numItems, reinterpret_cast<const uint4*>(pointersIn), numDrawables,
	                   reinterpret_cast<uint4*>(pointersOut));
	CADR_CHECK_HIP(hipGetLastError(), "finishCullKernel launch");
	return ok();
}
