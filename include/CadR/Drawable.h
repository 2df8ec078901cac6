// Geometry, MatrixList, Drawable and StateSet of the CadR facade.
//
// Same roles and semantics as the reference (src/CadR/Geometry.h,
// MatrixList.h, Drawable.h, StateSet.h):
//   - Geometry owns vertex, index and PrimitiveSet allocations, one handle each;
//   - MatrixList owns {numMatrices, capacity, pad[7]} + mat4[] (MatrixList.h:54-59);
//   - a Drawable binds Geometry + PrimitiveSet byte offset + MatrixList
//     (+ optional per-drawable data) into a 48-byte DrawableGpuData kept by
//     its StateSet (Drawable.cpp:35-78);
//   - a StateSet keeps its drawables in a vector; removal swaps the last
//     element into the hole (StateSet.cpp:29-47); child StateSets are visited
//     depth first, and a StateSet reachable through several parents is
//     visited once per path (StateSet.cpp:180-268).
// For the cull frame (Renderer::recordDrawableCulling, an extension: the
// reference draws every drawable) a Geometry also carries its local bounding
// sphere and LOD count (LOD k is the PrimitiveSet record at the drawable's
// offset + 8k), and the record's padding word holds the renderer's mesh
// index of the drawable's (Geometry, PrimitiveSet offset) pair.
#pragma once

#include "Export.h"

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <list>
#include <vector>

#include "DataStorage.h"

namespace CadR {

class Renderer;
class StateSet;

struct DrawableGpuData {
	uint64_t vertexDataHandle;
	uint64_t indexDataHandle;
	uint64_t matrixListHandle;
	uint64_t drawableDataHandle;
	uint64_t primitiveSetHandle;
	uint32_t primitiveSetOffset;
	uint32_t padding;
	DrawableGpuData() = default;
	constexpr DrawableGpuData(uint64_t v, uint64_t i, uint64_t m, uint64_t d, uint64_t p, uint32_t off)
	    : vertexDataHandle(v), indexDataHandle(i), matrixListHandle(m), drawableDataHandle(d),
	      primitiveSetHandle(p), primitiveSetOffset(off), padding(0) {}
};
static_assert(sizeof(DrawableGpuData) == 48, "DrawableGpuData is 48 bytes (Drawable.cpp:13)");

struct PrimitiveSet {
	uint32_t indexCount;
	uint32_t startIndex;
};

// Local bounding sphere (src/CadR/BoundingSphere.h:18-43: glm::vec3 centre and
// a float radius; the empty sphere has radius -inf).
struct BoundingSphere {
	float center[3];
	float radius;
	static constexpr BoundingSphere empty() { return BoundingSphere{{0.f, 0.f, 0.f}, -__builtin_huge_valf()}; }
	bool isEmpty() const { return radius == -__builtin_huge_valf(); }
};

class CADR_EXPORT Geometry {
public:
	explicit Geometry(Renderer& r);
	~Geometry() noexcept;
	Geometry(const Geometry&) = delete;
	Geometry& operator=(const Geometry&) = delete;

	// culling: local bounding sphere (empty by default: never visible) and the
	// number of LODs, PrimitiveSet records at offset, offset + 8, ... (1..4)
	void setBoundingSphere(const BoundingSphere& s) { _boundingSphere = s; }
	const BoundingSphere& boundingSphere() const { return _boundingSphere; }
	void setNumLods(uint32_t n);
	uint32_t numLods() const { return _numLods; }
	Renderer& renderer() const { return *_renderer; }

	DataAllocation& vertexDataAllocation() { return _vertices; }
	DataAllocation& indexDataAllocation() { return _indices; }
	DataAllocation& primitiveSetDataAllocation() { return _primitiveSets; }
	const DataAllocation& vertexDataAllocation() const { return _vertices; }
	const DataAllocation& indexDataAllocation() const { return _indices; }
	const DataAllocation& primitiveSetDataAllocation() const { return _primitiveSets; }

	void uploadVertexData(const void* ptr, size_t numBytes) { _vertices.upload(ptr, numBytes); }
	void uploadIndexData(const void* ptr, size_t numBytes) { _indices.upload(ptr, numBytes); }
	void uploadPrimitiveSetData(const void* ptr, size_t numBytes) { _primitiveSets.upload(ptr, numBytes); }
	void* createVertexStagingData(size_t numBytes) { return _vertices.alloc(numBytes); }
	void* createIndexStagingData(size_t numBytes) { return _indices.alloc(numBytes); }
	void* createPrimitiveSetStagingData(size_t numBytes) { return _primitiveSets.alloc(numBytes); }
	void freeData() { _vertices.free(); _indices.free(); _primitiveSets.free(); }

private:
	Renderer* _renderer;
	DataAllocation _vertices, _indices, _primitiveSets;
	BoundingSphere _boundingSphere = BoundingSphere::empty();
	uint32_t _numLods = 1;
};

class CADR_EXPORT MatrixList {
public:
	explicit MatrixList(Renderer& r);
	// column-major float mat4 values, 16 floats each
	void setMatrices(const float* matrices, size_t numMatrices);
	float* editNewContent(size_t numMatrices);
	uint64_t handle() const { return _matrixList.handle(); }
	uint64_t deviceAddress() const { return _matrixList.deviceAddress(); }
	// the matrices follow the 64-byte header: a device array of numMatrices()
	// column-major mat4 that a GPU transform pass may write directly
	uint64_t matricesDeviceAddress() const { return _matrixList.deviceAddress() + 64; }
	size_t numMatrices() const { return _numMatrices; }

private:
	Renderer* _renderer;
	DataAllocation _matrixList;
	size_t _numMatrices = 0;
};

class CADR_EXPORT Drawable {
public:
	Drawable() noexcept = default;
	Drawable(Geometry& geometry, uint32_t primitiveSetOffset, MatrixList& matrixList, StateSet& stateSet);
	Drawable(Geometry& geometry, uint32_t primitiveSetOffset, MatrixList& matrixList, DataAllocation& drawableData,
	         StateSet& stateSet);
	Drawable(const Drawable&) = delete;
	Drawable& operator=(const Drawable&) = delete;
	Drawable(Drawable&& other) noexcept;
	Drawable& operator=(Drawable&& rhs) noexcept;
	~Drawable() noexcept;

	void create(Geometry& geometry, uint32_t primitiveSetOffset, MatrixList& matrixList, StateSet& stateSet);
	void create(Geometry& geometry, uint32_t primitiveSetOffset, MatrixList& matrixList, DataAllocation& drawableData,
	            StateSet& stateSet);
	void destroy() noexcept;
	bool isValid() const { return _indexIntoStateSet != ~0u; }
	StateSet& stateSet() const { return *_stateSet; }
	uint32_t indexIntoStateSet() const { return _indexIntoStateSet; }

private:
	void createInternal(Geometry& g, uint32_t off, MatrixList& ml, DataAllocation* dd, StateSet& ss);
	const MatrixList* matrixList() const { return _matrixList; }
	StateSet* _stateSet = nullptr;
	MatrixList* _matrixList = nullptr;
	DataAllocation* _drawableData = nullptr;
	uint32_t _indexIntoStateSet = ~0u;
	friend class StateSet;
	friend class Renderer;
};

// Parent-child relation of StateSets (ParentChildList.h): one edge per
// parent -> child link, listed in the parent's childList and the child's
// parentList in append order.  A StateSet may have several parents and may
// be linked to the same child more than once.  childList.append(child) and
// parentList.append(parent) create an edge and return the iterator to it in
// that list; remove(iterator) deletes the edge from both lists; clear()
// removes every edge of the list; destroying a StateSet removes its edges.
// Dereferencing a childList iterator gives the child, a parentList iterator
// the parent.
struct StateSetEdge;

template <bool CHILDREN>
class CADR_EXPORT StateSetList {
	using Raw = std::list<StateSetEdge*>;

public:
	class iterator {
	public:
		using iterator_category = std::bidirectional_iterator_tag;
		using value_type = StateSet;
		using difference_type = std::ptrdiff_t;
		using pointer = StateSet*;
		using reference = StateSet&;
		iterator() = default;
		explicit iterator(Raw::iterator it) : _it(it) {}
		StateSet& operator*() const;
		StateSet* operator->() const { return &**this; }
		iterator& operator++() { ++_it; return *this; }
		iterator operator++(int) { iterator t = *this; ++_it; return t; }
		iterator& operator--() { --_it; return *this; }
		iterator operator--(int) { iterator t = *this; --_it; return t; }
		bool operator==(const iterator& o) const { return _it == o._it; }
		bool operator!=(const iterator& o) const { return _it != o._it; }
		Raw::iterator raw() const { return _it; }

	private:
		Raw::iterator _it;
	};
	using const_iterator = iterator;

	explicit StateSetList(StateSet& owner) noexcept : _owner(&owner) {}
	StateSetList(const StateSetList&) = delete;
	StateSetList& operator=(const StateSetList&) = delete;
	~StateSetList() { clear(); }

	iterator append(StateSet& other);
	void remove(iterator it) noexcept;
	void clear() noexcept { while (!_list.empty()) remove(begin()); }
	size_t size() const { return _list.size(); }
	bool empty() const { return _list.empty(); }
	StateSet& front() const { return *iterator(const_cast<Raw&>(_list).begin()); }
	StateSet& back() const { return *std::prev(end()); }
	iterator begin() const { return iterator(const_cast<Raw&>(_list).begin()); }
	iterator end() const { return iterator(const_cast<Raw&>(_list).end()); }

private:
	StateSet* _owner;
	Raw _list;
	friend class StateSetList<!CHILDREN>;
	friend class StateSet;
};

using ChildList = StateSetList<true>;
using ParentList = StateSetList<false>;

// One StateSet's share of the indirect buffers, as StateSet::recordToCommandBuffer
// issues it: commands [base, base + count) (StateSet.cpp:233-265).  After a
// cull frame, the StateSet's cull items (one per matrix of each drawable's
// MatrixList) are [itemBase, itemBase + itemCount) of each view's culled
// command range, of which the first `count` of the view's count buffer are
// written; without culling itemBase = base and itemCount = count.
struct DrawRange {
	StateSet* stateSet;
	size_t base;
	size_t count;
	size_t itemBase = 0;
	size_t itemCount = 0;
};

class CADR_EXPORT StateSet {
public:
	explicit StateSet(Renderer& r) noexcept : _renderer(&r) {}
	~StateSet() noexcept;
	StateSet(const StateSet&) = delete;
	StateSet& operator=(const StateSet&) = delete;

	Renderer& renderer() const { return *_renderer; }
	// parent-child relation (StateSet.h:69-70, ParentChildList.h)
	ChildList childList{*this};
	ParentList parentList{*this};

	// rendering (StateSet.cpp:180-268)
	size_t prepareRecording();
	void setForceRecording(bool v) { _forceRecording = v; }
	bool forceRecording() const { return _forceRecording; }
	void requestRecording() { _skipRecording = false; }
	bool skipRecording() const { return _skipRecording; }
	// appends this subtree's drawables (DFS) to the staging list and its draw ranges
	void recordToCommandBuffer(DrawableGpuData* staging, std::vector<DrawRange>& ranges, size_t& drawableCounter);

	// drawables
	void appendDrawable(Drawable& d, const DrawableGpuData& gpuData);
	static void removeDrawable(Drawable& d);
	void removeAllDrawables() noexcept;
	Drawable& getDrawable(size_t index) const { return *_drawablePtrList[index]; }
	size_t getNumDrawables() const { return _drawablePtrList.size(); }
	const std::vector<DrawableGpuData>& drawableDataList() const { return _drawableDataList; }

private:
	void appendDrawableInternal(Drawable& d, const DrawableGpuData& gpuData);
	void removeDrawableInternal(Drawable& d) noexcept;
	void markDirty(uint32_t index) noexcept;

	Renderer* _renderer;
	bool _skipRecording = true;
	bool _forceRecording = false;
	std::vector<DrawableGpuData> _drawableDataList;
	std::vector<Drawable*> _drawablePtrList;
	// records changed since the renderer last uploaded this StateSet: every
	// index whose record differs from the uploaded one at the same index is
	// listed (append, removal's move into the hole, in-place create), or
	// _allDirty is set (a new StateSet, or more changes than are worth
	// listing)
	std::vector<uint32_t> _dirty;
	bool _allDirty = true;
	friend class Drawable;
	friend class Renderer;
};

}  // namespace CadR
