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
#include <vector>

#include "internal.h"
#include "launch.h"
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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void propagateKernel(const float4* __restrict__ local,
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
	if constexpr (ROOT && !SPHERES) {
		// world = local: a straight copy, each instruction one contiguous KiB
		// (through registers and LDS the compiler kept the matrices on the
		// scratch stack)
#pragma unroll
		for (int k = 0; k < 4; k++) {
			const uint32_t q = uint32_t(k) * 64u + lane;
			if (first + q / 4u < count) storeF4Stream(world + (4 * first + q), loadF4Stream(local + (4 * first + q)));
		}
		return;
	}
	float4 L[4];
	if constexpr (ROOT) {
		loadMatrices(local, first, count, stage[wave], lane, L);
		storeMatrices(world, first, count, stage[wave], lane, L);
		if (live) spheres[i] = nodeSphere(L, meshes, numMeshes, meshIndex[i]);
		return;
	}
	// the parent index and the local block in flight together, then the
	// parent matrices while the block goes through LDS
	const uint32_t p = live ? parentIndex[i] : ~0u;
	float4 raw[4];
	issueMatrixLoads(local, first, count, lane, raw);
	const bool hasParent = p < parentCount;
	float4 P[4], praw[4];
	if (parentCount > 0) issueParentLoads(parentWorld, parentCount, p, lane, praw);
	stageMatrices(stage[wave], lane, raw, L);
	if (parentCount > 0) stageMatrices(stage[wave], lane, praw, P);  // the same LDS, after the local block
	float4 W[4];
	if (hasParent) {
		mat4Mul(P, L, W);
	} else {
#pragma unroll
		for (int c = 0; c < 4; c++) W[c] = L[c];
	}
	storeMatrices(world, first, count, stage[wave], lane, W);
	if constexpr (SPHERES) {
		if (live) spheres[i] = nodeSphere(W, meshes, numMeshes, meshIndex[i]);
	}
}

// 64 nodes of one level from `first` on, one per lane of the calling wave,
// computed as propagateKernel computes them; the root/sphere choice is the
// level's (uniform in the wave).
// The descriptor's addresses are used as global (address space 1) pointers:
// generic ones would compile to FLAT accesses, which count on the LDS counter
// too, so every wait for the LDS staging would also wait for the unit's
// matrix loads and stores in flight.
template <bool SPHERES>
__device__ __forceinline__ void propagateUnit(const cadrTransformLevel& L, uint64_t first, float4* stage,
                                              uint32_t lane)
{
	typedef const __attribute__((address_space(1))) uint32_t GlobalU32;
	const uint64_t count = L.count;
	const uint64_t i = first + lane;
	const bool live = i < count;
	float4 M[4], W[4];
	// the parent index and the local block in flight together, then the
	// parent matrix while the block goes through LDS
	const uint32_t p = (L.parentWorld && live) ? reinterpret_cast<GlobalU32*>(L.parentIndex)[i] : ~0u;
	float4 raw[4];
	issueMatrixLoads(reinterpret_cast<const GlobalF4*>(L.localMatrices), first, count, lane, raw);
	const bool hasParent = p < L.parentCount;
	const bool anyParent = L.parentWorld && L.parentCount > 0;  // the same in every lane
	float4 P[4], praw[4];
	if (anyParent)
		issueParentLoads(reinterpret_cast<const GlobalF4*>(L.parentWorld), L.parentCount, p, lane, praw);
	stageMatrices(stage, lane, raw, M);
	if (anyParent) stageMatrices(stage, lane, praw, P);  // the same LDS, after the local block
	if (hasParent) {
		mat4Mul(P, M, W);
	} else {
#pragma unroll
		for (int c = 0; c < 4; c++) W[c] = M[c];
	}
	storeMatrices(reinterpret_cast<GlobalF4*>(L.worldMatrices), first, count, stage, lane, W);
	if (SPHERES && L.spheres && live) {
		const uint32_t m = reinterpret_cast<GlobalU32*>(L.meshIndex)[i];
		float4 sphere = make_float4(0.f, 0.f, 0.f, -__builtin_inff());  // BoundingSphere::empty()
		if (m < L.numMeshes)  // the mesh's first 16 bytes: local centre and radius
			sphere = sphereTransform(W, loadF4(reinterpret_cast<const GlobalF4*>(L.meshes) + 4 * uint64_t(m)));
		storeF4(reinterpret_cast<GlobalF4*>(L.spheres) + i, sphere);
	}
}

// Many levels (of different trees) in one launch: the levels' nodes are cut
// into units of 64 (one wave each), numbered level after level; the grid's
// waves stride over the units.  Each unit is computed exactly as
// propagateKernel computes it, with the level's root/sphere choice made per
// unit (uniform in the wave).  Levels in one launch must not depend on each
// other; the caller orders dependent levels into successive launches.
constexpr uint32_t kMaxBatchLevels = 1024;

template <bool SPHERES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void propagateBatchKernel(const cadrTransformLevel* __restrict__ levels,
                                                            uint32_t numLevels)
{
	__shared__ float4 stage[4][256];
	__shared__ uint64_t firstUnit[kMaxBatchLevels + 1];
	const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
			l = __builtin_amdgcn_readfirstlane(lo);  // the same in every lane
		}
		if (l != cached) {  // the descriptor stays in scalar registers while the level lasts
			L = levels[l];
			cached = l;
		}
		propagateUnit<SPHERES>(L, (unit - firstUnit[l]) * 64u, stage[wave], lane);
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

// A transform plan's launch: one unit (level, first node) per wave, in the
// plan's order; workgroups are dispatched in index order, so the waves in
// flight at any time hold consecutive units of that order.
template <bool SPHERES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void propagatePlanKernel(const cadrTransformLevel* __restrict__ levels,
                                                           const uint2* __restrict__ units, uint64_t numUnits)
{
	__shared__ float4 stage[4][256];
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const uint64_t u = uint64_t(blockIdx.x) * 4u + wave;
	if (u >= numUnits) return;  // whole wave past the end (waves are independent)
	const uint2 unit = units[u];
	const cadrTransformLevel L = levels[unit.x];
	propagateUnit<SPHERES>(L, uint64_t(unit.y), stage[wave], lane);
}

// The sort key of each unit of a plan: the parent index of its first node
// (0 for a root level, which reads no parent).
__global__ __launch_bounds__(256) void planKeysKernel(const cadrTransformLevel* __restrict__ levels,
                                                      const uint2* __restrict__ units, uint64_t numUnits,
                                                      uint32_t* __restrict__ keys)
{
	const uint64_t u = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (u >= numUnits) return;
	const uint2 unit = units[u];
	const cadrTransformLevel& L = levels[unit.x];
	keys[u] = L.parentWorld ? reinterpret_cast<const uint32_t*>(L.parentIndex)[unit.y] : 0u;
}

}  // namespace cadr

struct cadrTransformPlan {
	int device = 0;
	cadrTransformLevel* levels = nullptr;  // device copy of the descriptors
	uint2* units = nullptr;                // (level, first node), in launch order
	uint64_t numUnits = 0;
	bool spheres = false;                  // some level writes world spheres
};

using namespace cadr;

CADR_API int cadrPropagateTransforms(const cadrTransformLevel* lv, cadrStream stream)
{
	CADR_REQUIRE(lv, "cadrPropagateTransforms: level is NULL");
	if (lv->count == 0) CADR_RETURN_NO_LAUNCH(stream);
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
	auto kernel = root ? (spheres ? propagateKernel<true, true> : propagateKernel<true, false>)
	                   : (spheres ? propagateKernel<false, true> : propagateKernel<false, false>);
	launchKernel(kernel, g, b, s, local, parent, pworld, lv->parentCount, world, lv->count, mi, meshes, lv->numMeshes,
	             sph);
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
	bool spheres = false;
	for (uint32_t l = 0; l < numLevels; l++) {
		if (hostLevels[l].count == 0) continue;
		const int rc = checkLevel(&hostLevels[l], "cadrPropagateTransformsBatch");
		if (rc != CADR_SUCCESS) return rc;
		units += (hostLevels[l].count + 63u) / 64u;
		spheres = spheres || hostLevels[l].spheres != 0;
	}
	if (units == 0) CADR_RETURN_NO_LAUNCH(stream);
	// without world spheres the kernel holds fewer registers: 8 waves per SIMD instead of 5
	const uint64_t blocks = std::min<uint64_t>((units + 3) / 4, uint64_t(computeUnitCount()) * 16u);
	launchKernel(spheres ? propagateBatchKernel<true> : propagateBatchKernel<false>, dim3(uint32_t(blocks)), dim3(256),
	             hipStream_t(stream), reinterpret_cast<const cadrTransformLevel*>(deviceLevels), numLevels);
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

CADR_API int cadrTransformPlanCreate(const cadrTransformLevel* hostLevels, uint32_t numLevels, cadrStream stream,
                                     cadrTransformPlan** plan)
{
	CADR_REQUIRE(plan, "cadrTransformPlanCreate: plan is NULL");
	*plan = nullptr;
	CADR_REQUIRE(numLevels == 0 || hostLevels, "cadrTransformPlanCreate: hostLevels is NULL");
	std::vector<uint2> units;
	for (uint32_t l = 0; l < numLevels; l++) {
		if (hostLevels[l].count == 0) continue;
		const int rc = checkLevel(&hostLevels[l], "cadrTransformPlanCreate");
		if (rc != CADR_SUCCESS) return rc;
		CADR_REQUIRE(hostLevels[l].count <= 0xffffffffull, "cadrTransformPlanCreate: a level holds 2^32 nodes or more");
		for (uint64_t f = 0; f < hostLevels[l].count; f += 64) units.push_back(make_uint2(l, uint32_t(f)));
	}
	auto* p = new cadrTransformPlan;
	for (uint32_t l = 0; l < numLevels; l++) p->spheres = p->spheres || (hostLevels[l].count && hostLevels[l].spheres);
	CADR_CHECK_HIP(hipGetDevice(&p->device), "hipGetDevice");
	p->numUnits = units.size();
	hipStream_t s = hipStream_t(stream);
	uint32_t* dKeys = nullptr;
	auto fail = [&](hipError_t e, const char* what) {
		(void)hipFree(dKeys);
		cadrTransformPlanDestroy(p);
		return setHipError(e, what);
	};
	hipError_t e = hipMalloc(&p->levels, std::max<size_t>(numLevels, 1) * sizeof(cadrTransformLevel));
	if (e != hipSuccess) return fail(e, "hipMalloc(plan levels)");
	e = hipMalloc(&p->units, std::max<size_t>(units.size(), 1) * sizeof(uint2));
	if (e != hipSuccess) return fail(e, "hipMalloc(plan units)");
	if (!units.empty()) {
		e = hipMalloc(&dKeys, units.size() * sizeof(uint32_t));
		if (e != hipSuccess) return fail(e, "hipMalloc(plan keys)");
		e = hipMemcpyAsync(p->levels, hostLevels, numLevels * sizeof(cadrTransformLevel), hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = hipMemcpyAsync(p->units, units.data(), units.size() * sizeof(uint2), hipMemcpyHostToDevice, s);
		if (e != hipSuccess) return fail(e, "hipMemcpyAsync(plan)");
		hipLaunchKernelGGL(planKeysKernel, dim3(uint32_t((units.size() + 255) / 256)), dim3(256), 0, s, p->levels,
		                   p->units, uint64_t(units.size()), dKeys);
		e = hipGetLastError();
		if (e != hipSuccess) return fail(e, "planKeysKernel launch");
		std::vector<uint32_t> keys(units.size());
		e = hipMemcpyAsync(keys.data(), dKeys, keys.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
		if (e == hipSuccess) e = hipStreamSynchronize(s);
		if (e != hipSuccess) return fail(e, "hipMemcpyAsync(plan keys)");
		// order: by the parent array, then by the parent of the unit's first
		// node; equal keys keep their level-by-level order
		std::vector<uint32_t> order(units.size());
		for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
		std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
			const uint64_t pa = hostLevels[units[a].x].parentWorld, pb = hostLevels[units[b].x].parentWorld;
			if (pa != pb) return pa < pb;
			return keys[a] < keys[b];
		});
		std::vector<uint2> sorted(units.size());
		for (size_t i = 0; i < order.size(); i++) sorted[i] = units[order[i]];
		e = hipMemcpy(p->units, sorted.data(), sorted.size() * sizeof(uint2), hipMemcpyHostToDevice);
		if (e != hipSuccess) return fail(e, "hipMemcpy(plan units)");
		(void)hipFree(dKeys);
		dKeys = nullptr;
	}
	*plan = p;
	return ok();
}

CADR_API int cadrPropagateTransformsPlanned(const cadrTransformPlan* plan, cadrStream stream)
{
	CADR_REQUIRE(plan, "cadrPropagateTransformsPlanned: plan is NULL");
	if (plan->numUnits == 0) CADR_RETURN_NO_LAUNCH(stream);
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == plan->device, "cadrPropagateTransformsPlanned: the plan belongs to another device");
	const uint64_t blocks = (plan->numUnits + 3) / 4;
	CADR_REQUIRE(blocks < (1ull << 31), "cadrPropagateTransformsPlanned: plan too large");
	launchKernel(plan->spheres ? propagatePlanKernel<true> : propagatePlanKernel<false>, dim3(uint32_t(blocks)),
	             dim3(256), hipStream_t(stream), plan->levels, plan->units, plan->numUnits);
	CADR_CHECK_HIP(hipGetLastError(), "propagatePlanKernel launch");
	return ok();
}

CADR_API void cadrTransformPlanDestroy(cadrTransformPlan* plan)
{
	if (!plan) return;
	(void)hipFree(plan->levels);
	(void)hipFree(plan->units);
	delete plan;
}
