// Transform propagation by tree level and world bounding spheres.
//
// The reference propagates a node hierarchy on the CPU by recursion,
// m = parentMatrix * node.matrix (examples/gltfReader/main.cpp:1538-1576), and
// transforms a BoundingSphere by a mat4 on the CPU (src/CadR/BoundingSphere.h:70-87).
// Here a tree is stored level by level (all nodes of depth d contiguous, each
// with the uint32 index of its parent in level d-1), so one launch per level
// computes every node of that level in parallel; levels run in order on one
// stream.  The leaf level can also write the drawables' world spheres, so
// the sphere transform costs no extra pass over the world matrices.
//
// Memory: per node 64 B local read + 64 B world write (+4 B parent index);
// parents are few and shared by many children, so their reads hit L2.  The
// matrices are stored as the reference stores them (column-major glm::mat4,
// 64 B), which keeps the world array usable as MatrixList payload; a wave's
// 64 matrices are staged through LDS so that every global load and store
// instruction moves one contiguous KiB.
#include <algorithm>
#include <string>

#include "internal.h"
#include "fp32.h"
#include "matrix_io.h"

namespace cadr {

// World sphere of a node: the mesh's local sphere transformed by W, or the
// empty sphere of BoundingSphere::empty() (BoundingSphere.h:29-31, radius
// -inf, never visible) for a mesh index outside the table.
__device__ __forceinline__ float4 nodeSphere(const float4 W[4], const cadrMesh* __restrict__ meshes,
                                             uint32_t numMeshes, uint32_t m)
{
	if (m >= numMeshes) return make_float4(0.f, 0.f, 0.f, -__builtin_inff());
	const float4 local = *reinterpret_cast<const float4*>(meshes + m);  // center.xyz, radius
	return sphereTransform(W, local);
}

template <bool ROOT, bool SPHERES>
__global__ __launch_bounds__(256) void propagateKernel(const float4* __restrict__ local,
                                                       const uint32_t* __restrict__ parentIndex,
                                                       const float4* __restrict__ parentWorld,
                                                       uint32_t parentCount, float4* __restrict__ world,
                                                       uint64_t count,
                                                       const uint32_t* __restrict__ meshIndex,
                                                       const cadrMesh* __restrict__ meshes, uint32_t numMeshes,
                                                       float4* __restrict__ spheres)
{
	__shared__ float4 stage[4][256];
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	const uint64_t first = uint64_t(blockIdx.x) * 256u + wave * 64u;  // the wave's first node
	if (first >= count) return;  // whole wave past the end (waves are independent)
	const uint64_t i = first + lane;
	const bool live = i < count;
	float4 L[4];
	loadMatrices(local, first, count, stage[wave], lane, L);
	float4 W[4];
	if constexpr (ROOT) {
#pragma unroll
		for (int c = 0; c < 4; c++) W[c] = L[c];
	} else {
		const uint32_t p = live ? parentIndex[i] : ~0u;
		if (p < parentCount) {
			float4 P[4];
#pragma unroll
			for (int c = 0; c < 4; c++) P[c] = parentWorld[4 * uint64_t(p) + c];
			mat4Mul(P, L, W);
		} else {
#pragma unroll
			for (int c = 0; c < 4; c++) W[c] = L[c];
		}
	}
	storeMatrices(world, first, count, stage[wave], lane, W);
	if constexpr (SPHERES) {
		if (live) spheres[i] = nodeSphere(W, meshes, numMeshes, meshIndex[i]);
	}
}

// Many levels (of different trees) in one launch: the levels' nodes are cut
// into units of 64 (one wave each), numbered level after level; the grid's
// waves stride over the units.  Each unit is computed exactly as
// propagateKernel computes it, with the level's root/sphere choice made per
// unit (uniform in the wave).  Levels in one launch must not depend on each
// other; the caller orders dependent levels into successive launches.
constexpr uint32_t kMaxBatchLevels = 1024;

__global__ __launch_bounds__(256) void propagateBatchKernel(const cadrTransformLevel* __restrict__ levels,
                                                            uint32_t numLevels)
{
	__shared__ float4 stage[4][256];
	__shared__ uint64_t firstUnit[kMaxBatchLevels + 1];
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	// firstUnit[l] = units of levels 0..l-1: loaded in parallel, then an
	// in-place Hillis-Steele scan over the (at most 1024) levels
	for (uint32_t i = threadIdx.x; i < numLevels; i += 256u) firstUnit[i + 1] = (levels[i].count + 63u) / 64u;
	if (threadIdx.x == 0) firstUnit[0] = 0;
	__syncthreads();
	for (uint32_t off = 1; off < numLevels; off <<= 1) {
		uint64_t add[kMaxBatchLevels / 256];
#pragma unroll
		for (uint32_t k = 0; k < kMaxBatchLevels / 256; k++) {
			const uint32_t i = threadIdx.x + 256u * k;
			add[k] = (i < numLevels && i >= off) ? firstUnit[i + 1 - off] : 0u;
		}
		__syncthreads();
#pragma unroll
		for (uint32_t k = 0; k < kMaxBatchLevels / 256; k++) {
			const uint32_t i = threadIdx.x + 256u * k;
			if (i < numLevels) firstUnit[i + 1] += add[k];
		}
		__syncthreads();
	}
	const uint64_t totalUnits = firstUnit[numLevels];
	const uint64_t waves = uint64_t(gridDim.x) * 4u;
	uint32_t l = 0, cached = ~0u;
	cadrTransformLevel L{};
	for (uint64_t unit = uint64_t(blockIdx.x) * 4u + wave; unit < totalUnits; unit += waves) {
		// the level holding this unit: units only grow, so search upwards from
		// the last one (a grid stride can skip many small levels)
		if (firstUnit[l + 1] <= unit) {
			uint32_t lo = l + 1, hi = numLevels - 1;
			while (lo < hi) {
				const uint32_t mid = (lo + hi + 1) / 2;
				if (firstUnit[mid] <= unit) lo = mid; else hi = mid - 1;
			}
			l = lo;
		}
		if (l != cached) {  // the descriptor stays in scalar registers while the level lasts
			L = levels[l];
			cached = l;
		}
		const uint64_t count = L.count;
		const uint64_t first = (unit - firstUnit[l]) * 64u;
		const uint64_t i = first + lane;
		const bool live = i < count;
		const float4* local = reinterpret_cast<const float4*>(L.localMatrices);
		float4 M[4], W[4];
		loadMatrices(local, first, count, stage[wave], lane, M);
		const uint32_t p = (L.parentWorld && live) ? reinterpret_cast<const uint32_t*>(L.parentIndex)[i] : ~0u;
		if (p < L.parentCount) {
			const float4* pw = reinterpret_cast<const float4*>(L.parentWorld);
			float4 P[4];
#pragma unroll
			for (int c = 0; c < 4; c++) P[c] = pw[4 * uint64_t(p) + c];
			mat4Mul(P, M, W);
		} else {
#pragma unroll
			for (int c = 0; c < 4; c++) W[c] = M[c];
		}
		storeMatrices(reinterpret_cast<float4*>(L.worldMatrices), first, count, stage[wave], lane, W);
		if (L.spheres && live)
			reinterpret_cast<float4*>(L.spheres)[i] =
			    nodeSphere(W, reinterpret_cast<const cadrMesh*>(L.meshes), L.numMeshes,
			               reinterpret_cast<const uint32_t*>(L.meshIndex)[i]);
	}
}

__global__ __launch_bounds__(256) void worldSpheresKernel(const float4* __restrict__ world,
                                                          const uint32_t* __restrict__ meshIndex,
                                                          const cadrMesh* __restrict__ meshes, uint32_t numMeshes,
                                                          uint64_t count, float4* __restrict__ spheres)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i >= count) return;
	float4 W[4];
#pragma unroll
	for (int c = 0; c < 4; c++) W[c] = world[4 * i + c];
	spheres[i] = nodeSphere(W, meshes, numMeshes, meshIndex[i]);
}

}  // namespace cadr

using namespace cadr;

CADR_API int cadrPropagateTransforms(const cadrTransformLevel* lv, cadrStream stream)
{
	CADR_REQUIRE(lv, "cadrPropagateTransforms: level is NULL");
	if (lv->count == 0) return ok();
	CADR_REQUIRE(lv->localMatrices && lv->worldMatrices, "cadrPropagateTransforms: null matrix array");
	CADR_REQUIRE((lv->localMatrices % 16) == 0 && (lv->worldMatrices % 16) == 0 && (lv->parentWorld % 16) == 0,
	             "cadrPropagateTransforms: matrix arrays must be 16-byte aligned");
	const bool root = lv->parentWorld == 0;
	CADR_REQUIRE(root || lv->parentIndex, "cadrPropagateTransforms: parentIndex is NULL for a non-root level");
	const bool spheres = lv->spheres != 0;
	CADR_REQUIRE(!spheres || (lv->meshIndex && lv->meshes && (lv->spheres % 16) == 0 && (lv->meshes % 16) == 0),
	             "cadrPropagateTransforms: spheres need meshIndex and 16-byte aligned meshes/spheres");
	const uint64_t blocks = (lv->count + 255) / 256;
	CADR_REQUIRE(blocks < (1ull << 31), "cadrPropagateTransforms: level too large");
	auto* local = reinterpret_cast<const float4*>(lv->localMatrices);
	auto* parent = reinterpret_cast<const uint32_t*>(lv->parentIndex);
	auto* pworld = reinterpret_cast<const float4*>(lv->parentWorld);
	auto* world = reinterpret_cast<float4*>(lv->worldMatrices);
	auto* mi = reinterpret_cast<const uint32_t*>(lv->meshIndex);
	auto* meshes = reinterpret_cast<const cadrMesh*>(lv->meshes);
	auto* sph = reinterpret_cast<float4*>(lv->spheres);
	hipStream_t s = hipStream_t(stream);
	const dim3 g{uint32_t(blocks)}, b{256};
	if (root && spheres) hipLaunchKernelGGL((propagateKernel<true, true>), g, b, 0, s, local, parent, pworld, lv->parentCount, world, lv->count, mi, meshes, lv->numMeshes, sph);
	else if (root) hipLaunchKernelGGL((propagateKernel<true, false>), g, b, 0, s, local, parent, pworld, lv->parentCount, world, lv->count, mi, meshes, lv->numMeshes, sph);
	else if (spheres) hipLaunchKernelGGL((propagateKernel<false, true>), g, b, 0, s, local, parent, pworld, lv->parentCount, world, lv->count, mi, meshes, lv->numMeshes, sph);
	else hipLaunchKernelGGL((propagateKernel<false, false>), g, b, 0, s, local, parent, pworld, lv->parentCount, world, lv->count, mi, meshes, lv->numMeshes, sph);
	CADR_CHECK_HIP(hipGetLastError(), "propagateKernel launch");
	return ok();
}

namespace {
int checkLevel(const cadrTransformLevel* lv, const char* who)
{
	std::string w(who);
	CADR_REQUIRE(lv->localMatrices && lv->worldMatrices, (w + ": null matrix array").c_str());
	CADR_REQUIRE((lv->localMatrices % 16) == 0 && (lv->worldMatrices % 16) == 0 && (lv->parentWorld % 16) == 0,
	             (w + ": matrix arrays must be 16-byte aligned").c_str());
	CADR_REQUIRE(lv->parentWorld == 0 || lv->parentIndex, (w + ": parentIndex is NULL for a non-root level").c_str());
	CADR_REQUIRE(lv->spheres == 0 || (lv->meshIndex && lv->meshes && (lv->spheres % 16) == 0 && (lv->meshes % 16) == 0),
	             (w + ": spheres need meshIndex and 16-byte aligned meshes/spheres").c_str());
	return CADR_SUCCESS;
}
}  // namespace

CADR_API int cadrPropagateTransformsBatch(uint64_t deviceLevels, const cadrTransformLevel* hostLevels,
                                          uint32_t numLevels, cadrStream stream)
{
	CADR_REQUIRE(numLevels == 0 || (hostLevels && deviceLevels), "cadrPropagateTransformsBatch: null argument");
	CADR_REQUIRE(numLevels <= kMaxBatchLevels, "cadrPropagateTransformsBatch: more than 1024 levels");
	CADR_REQUIRE(deviceLevels % 8 == 0, "cadrPropagateTransformsBatch: misaligned level array");
	uint64_t units = 0;
	for (uint32_t l = 0; l < numLevels; l++) {
		if (hostLevels[l].count == 0) continue;
		const int rc = checkLevel(&hostLevels[l], "cadrPropagateTransformsBatch");
		if (rc != CADR_SUCCESS) return rc;
		units += (hostLevels[l].count + 63u) / 64u;
	}
	if (units == 0) return ok();
	const uint64_t blocks = std::min<uint64_t>((units + 3) / 4, uint64_t(computeUnitCount()) * 16u);
	hipLaunchKernelGGL(propagateBatchKernel, dim3(uint32_t(blocks)), dim3(256), 0, hipStream_t(stream),
	                   reinterpret_cast<const cadrTransformLevel*>(deviceLevels), numLevels);
	CADR_CHECK_HIP(hipGetLastError(), "propagateBatchKernel launch");
	return ok();
}

CADR_API int cadrComputeWorldSpheres(uint64_t worldMatrices, uint64_t meshIndex, uint64_t meshes,
                                     uint32_t numMeshes, uint64_t count, uint64_t spheres, cadrStream stream)
{
	if (count == 0) return ok();
	CADR_REQUIRE(worldMatrices && meshIndex && meshes && spheres, "cadrComputeWorldSpheres: null address");
	CADR_REQUIRE((worldMatrices % 16) == 0 && (meshes % 16) == 0 && (spheres % 16) == 0,
	             "cadrComputeWorldSpheres: arrays must be 16-byte aligned");
	const uint64_t blocks = (count + 255) / 256;
	CADR_REQUIRE(blocks < (1ull << 31), "cadrComputeWorldSpheres: count too large");
	hipLaunchKernelGGL(worldSpheresKernel, dim3(uint32_t(blocks)), dim3(256), 0, hipStream_t(stream),
	                   reinterpret_cast<const float4*>(worldMatrices), reinterpret_cast<const uint32_t*>(meshIndex),
	                   reinterpret_cast<const cadrMesh*>(meshes), numMeshes, count, reinterpret_cast<float4*>(spheres));
	CADR_CHECK_HIP(hipGetLastError(), "worldSpheresKernel launch");
	return ok();
}
