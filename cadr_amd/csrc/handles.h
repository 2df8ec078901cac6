// Handle-table lookup and the DrawableGpuData record, shared by the gather
// (processDrawables.comp) and the facade's cull-input kernels.
#pragma once

#include "internal.h"

namespace cadr {

struct alignas(16) DrawableRecord {
	uint64_t vertexDataHandle;
	uint64_t indexDataHandle;
	uint64_t matrixListHandle;
	uint64_t drawableDataHandle;
	uint64_t primitiveSetHandle;
	uint32_t primitiveSetOffset;
	uint32_t padding;
};
static_assert(sizeof(DrawableRecord) == 48, "DrawableGpuData is 48 bytes (Drawable.cpp:13)");

// Handle -> device address through 1, 2 or 3 tables of 2048 uint64 entries
// (processDrawables.comp:77-89: level 1 table[h], level 2 table[h>>11][h&0x7ff],
// level 3 table[h>>22][(h>>11)&0x7ff][h&0x7ff], indices truncated to 32 bits
// as uint(handle) does in GLSL).  An index that falls outside a table, or a
// table pointer that is 0, resolves to 0 instead of reading unowned memory.
template <int LEVEL>
__device__ __forceinline__ uint64_t lookupHandle(const uint64_t* __restrict__ root, uint64_t handle)
{
	const uint32_t h = uint32_t(handle);
	if constexpr (LEVEL == 1) {
		if (h >= CADR_HANDLES_PER_TABLE) return 0;
		return root[h];
	} else if constexpr (LEVEL == 2) {
		const uint32_t i1 = uint32_t(handle >> CADR_HANDLE_LEVEL_SHIFT);
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		const uint64_t* t2 = reinterpret_cast<const uint64_t*>(root[i1]);
		if (!t2) return 0;
		return t2[h & CADR_HANDLE_LEVEL_MASK];
	} else {
		const uint32_t i1 = uint32_t(handle >> (2 * CADR_HANDLE_LEVEL_SHIFT));
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		const uint64_t* t2 = reinterpret_cast<const uint64_t*>(root[i1]);
		if (!t2) return 0;
		const uint64_t* t3 = reinterpret_cast<const uint64_t*>(
		    t2[uint32_t(handle >> CADR_HANDLE_LEVEL_SHIFT) & CADR_HANDLE_LEVEL_MASK]);
		if (!t3) return 0;
		return t3[h & CADR_HANDLE_LEVEL_MASK];
	}
}

}  // namespace cadr
