// Wave-wide movement of column-major float mat4 arrays between HBM and
// registers, shared by the propagation kernels and the fused
// propagation + cull kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace cadr {

// 16-byte element access through any pointer type (generic or an explicit
// address space), component by component so that address-space-qualified
// pointers need no float4 copy constructor; the loads and stores merge into
// one 16-byte instruction.
template <class P>
__device__ __forceinline__ float4 loadF4(P p)
{
	return make_float4(p->x, p->y, p->z, p->w);
}

template <class P>
__device__ __forceinline__ void storeF4(P p, float4 v)
{
	p->x = v.x;
	p->y = v.y;
	p->z = v.z;
	p->w = v.w;
}

// Streaming (nontemporal) 16-byte access for data touched once per pass:
// the local matrices a propagation reads and the world matrices it writes.
// Marked so, they do not displace what the caches should keep (parent world
// matrices, the mesh table, command lines the L2 is still merging): the
// full10m frame's fused leaf + cull launch ran 0.295 -> 0.266 ms with these.
// One overload per address space (generic pointers in the propagation
// kernels, global ones in the cull kernels), so the access stays a global
// instruction.
typedef float StreamF4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) float4 GlobalF4;
typedef __attribute__((address_space(1))) StreamF4 GlobalStreamF4;

__device__ __forceinline__ float4 loadF4Stream(const float4* p)
{
	const StreamF4 x = __builtin_nontemporal_load(reinterpret_cast<const StreamF4*>(p));
	return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ float4 loadF4Stream(const GlobalF4* p)
{
	const StreamF4 x = __builtin_nontemporal_load((const GlobalStreamF4*)p);
	return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void storeF4Stream(float4* p, float4 v)
{
	const StreamF4 x = {v.x, v.y, v.z, v.w};
	__builtin_nontemporal_store(x, reinterpret_cast<StreamF4*>(p));
}
__device__ __forceinline__ void storeF4Stream(GlobalF4* p, float4 v)
{
	const StreamF4 x = {v.x, v.y, v.z, v.w};
	__builtin_nontemporal_store(x, (GlobalStreamF4*)p);
}

// LDS slot of 16-byte unit q (matrix q/4, column q%4) of a wave's staged
// block.  Columns are rotated within their matrix by ((m >> 1) ^ (m >> 3)) for
// matrix m, so that both the lane-contiguous accesses (lane l: unit 64k + l)
// and the per-matrix ones (lane l: units 4l..4l+3, one column at a time)
// touch distinct 16-byte bank groups in every 8 and every 16 consecutive
// lanes: no LDS bank conflicts either way (the plain layout puts every
// second lane of the per-matrix access in the same banks).
__device__ __forceinline__ uint32_t stageSlot(uint32_t q)
{
	const uint32_t m = q >> 2;
	return (m << 2) | (((q & 3u) + ((m >> 1) ^ (m >> 3))) & 3u);
}

// A wave's 64 matrices (4 KiB, contiguous in memory) move between HBM and
// registers through LDS: four 16-byte loads per lane at lane stride 16 B (each
// instruction one contiguous KiB), then each lane reads its own 64 bytes from
// LDS; stores go the same way in reverse, both as streaming accesses.  Matrix
// `first` is lane 0's; count >= 1.  The four loads are unconditional, so all
// of them are in flight at once (a guarded load per instruction made the
// compiler wait for each before issuing the next): lanes whose matrix index
// is at or past `count` read matrix count - 1 again, and nothing of theirs is
// stored.
template <class Src>
__device__ __forceinline__ void issueMatrixLoads(Src src, uint64_t first, uint64_t count, uint32_t lane,
                                                 float4 raw[4])
{
#pragma unroll
	for (int k = 0; k < 4; k++) {
		const uint32_t q = uint32_t(k) * 64u + lane;      // float4 index inside the wave's 4 KiB
		const uint64_t m = min(first + q / 4u, count - 1u);
		raw[k] = loadF4Stream(src + (4 * m + (q & 3u)));
	}
}

// The loaded block (issueMatrixLoads) redistributed through LDS: lane l gets
// matrix first + l.
__device__ __forceinline__ void stageMatrices(float4* stage, uint32_t lane, const float4 raw[4], float4 M[4])
{
#pragma unroll
	for (int k = 0; k < 4; k++) stage[stageSlot(uint32_t(k) * 64u + lane)] = raw[k];
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
	for (int c = 0; c < 4; c++) M[c] = stage[stageSlot(4 * lane + c)];
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The parent world matrices of a wave's 64 nodes (lane l's parent index p),
// read whole: lane l loads column l % 4 of the parent of lane 16k + l / 4 for
// k = 0..3, so each instruction reads 16 parents (one or two 128-byte lines
// each) instead of one column of 64 parents scattered over 64 lines; the
// block then goes through LDS (stageMatrices), which hands lane l its own
// parent.  Indices at or past parentCount read parent 0 (parentCount >= 1),
// and the caller ignores those lanes' matrices.
template <class Src>
__device__ __forceinline__ void issueParentLoads(Src parentWorld, uint32_t parentCount, uint32_t p, uint32_t lane,
                                                 float4 raw[4])
{
#pragma unroll
	for (int k = 0; k < 4; k++) {
		const uint32_t q = uint32_t(__shfl(int(p), int(16u * uint32_t(k) + (lane >> 2))));
		raw[k] = loadF4(parentWorld + (4 * uint64_t(q < parentCount ? q : 0u) + (lane & 3u)));
	}
}

template <class Src>
__device__ __forceinline__ void loadMatrices(Src src, uint64_t first, uint64_t count, float4* stage, uint32_t lane,
                                             float4 M[4])
{
	float4 raw[4];
	issueMatrixLoads(src, first, count, lane, raw);
	stageMatrices(stage, lane, raw, M);
}

template <class Dst>
__device__ __forceinline__ void storeMatrices(Dst dst, uint64_t first, uint64_t count, float4* stage, uint32_t lane,
                                              const float4 M[4])
{
#pragma unroll
	for (int c = 0; c < 4; c++) stage[stageSlot(4 * lane + c)] = M[c];
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	float4 out[4];  // all four LDS reads first, then the stores
#pragma unroll
	for (int k = 0; k < 4; k++) out[k] = stage[stageSlot(uint32_t(k) * 64u + lane)];
#pragma unroll
	for (int k = 0; k < 4; k++) {
		const uint32_t q = uint32_t(k) * 64u + lane;
		if (first + q / 4u < count) storeF4Stream(dst + (4 * first + q), out[k]);
	}
}

}  // namespace cadr
