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
// Addresses read from memory are used as global (address space 1)
// pointers: as generic ones they would compile to FLAT loads.
template <class T>
using GlobalConst = const __attribute__((address_space(1))) T*;

template <class T>
using GlobalMut = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ GlobalConst<T> globalAt(uint64_t address)
{
	return reinterpret_cast<GlobalConst<T>>(address);
}
template <class T>
__device__ __forceinline__ GlobalConst<T> globalAt(const T* p)
{
	return reinterpret_cast<GlobalConst<T>>(reinterpret_cast<uint64_t>(p));
}
template <class T>
__device__ __forceinline__ GlobalMut<T> globalMutAt(T* p)
{
	return reinterpret_cast<GlobalMut<T>>(reinterpret_cast<uint64_t>(p));
}

// 16-byte load through an address-space-qualified pointer, component by
// component (the vector type has no copy from another address space); the
// four loads merge into one instruction.
__device__ __forceinline__ uint4 loadU4(GlobalConst<uint4> p)
{
	return make_uint4(p->x, p->y, p->z, p->w);
}

template <int LEVEL>
__device__ __forceinline__ uint64_t lookupHandle(const uint64_t* __restrict__ rootTable, uint64_t handle)
{
	const GlobalConst<uint64_t> root = globalAt<uint64_t>(reinterpret_cast<uint64_t>(rootTable));
	const uint32_t h = uint32_t(handle);
	if constexpr (LEVEL == 1) {
		if (h >= CADR_HANDLES_PER_TABLE) return 0;
		return root[h];
	} else if constexpr (LEVEL == 2) {
		const uint32_t i1 = uint32_t(handle >> CADR_HANDLE_LEVEL_SHIFT);
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		const uint64_t t2 = root[i1];
		if (!t2) return 0;
		return globalAt<uint64_t>(t2)[h & CADR_HANDLE_LEVEL_MASK];
	} else {
		const uint32_t i1 = uint32_t(handle >> (2 * CADR_HANDLE_LEVEL_SHIFT));
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		const uint64_t t2 = root[i1];
		if (!t2) return 0;
		const uint64_t t3 = globalAt<uint64_t>(t2)[uint32_t(handle >> CADR_HANDLE_LEVEL_SHIFT) & CADR_HANDLE_LEVEL_MASK];
		if (!t3) return 0;
		return globalAt<uint64_t>(t3)[h & CADR_HANDLE_LEVEL_MASK];
	}
}

}  // namespace cadr
