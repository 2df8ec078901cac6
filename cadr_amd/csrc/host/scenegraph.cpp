// Geometry, MatrixList, Drawable and StateSet of the CadR facade.
#include <algorithm>
#include <iterator>
#include <list>
#include <cstring>

#include "../../../include/CadR/Drawable.h"
#include "../../../include/CadR/Exceptions.h"
#include "../../../include/CadR/Renderer.h"

namespace CadR {

Geometry::Geometry(Renderer& r)
    : _renderer(&r), _vertices(r.dataStorage()), _indices(r.dataStorage()), _primitiveSets(r.dataStorage())
{
}

// the renderer's mesh entries of this Geometry stop existing (never visible)
Geometry::~Geometry() noexcept { _renderer->forgetGeometry(this); }

void Geometry::setNumLods(uint32_t n)
{
	if (n < 1 || n > 4) throw LogicError("CadR::Geometry::setNumLods(): the LOD count must be 1 to 4");
	_numLods = n;
}

// ------------------------------------------------------------ MatrixList

MatrixList::MatrixList(Renderer& r) : _renderer(&r), _matrixList(r.dataStorage()) {}

// Header {numMatrices, capacity = numMatrices, 0 x 7} then the matrices
// (MatrixList.h:54-59: initZeroField writes both counts and zeroes the rest).
float* MatrixList::editNewContent(size_t numMatrices)
{
	auto* p = static_cast<uint8_t*>(_matrixList.alloc(64 + 64 * numMatrices));
	if (numMatrices != _numMatrices) {
		_numMatrices = numMatrices;
		_renderer->sceneChanged();  // the cull items of its drawables change
	}
	uint32_t hdr[16] = {};
	hdr[0] = uint32_t(numMatrices);
	hdr[1] = uint32_t(numMatrices);
	std::memcpy(p, hdr, 64);
	return reinterpret_cast<float*>(p + 64);
}

void MatrixList::setMatrices(const float* matrices, size_t numMatrices)
{
	float* m = editNewContent(numMatrices);
	if (numMatrices) std::memcpy(m, matrices, 64 * numMatrices);
}

// -------------------------------------------------------------- Drawable

// The padding word carries the renderer's mesh index of (g, off) for the cull
// frame; the gather does not read it.
static DrawableGpuData gpuDataOf(Geometry& g, uint32_t off, MatrixList& ml, DataAllocation* dd)
{
	DrawableGpuData d(g.vertexDataAllocation().handle(), g.indexDataAllocation().handle(), ml.handle(),
	                  dd ? dd->handle() : 0, g.primitiveSetDataAllocation().handle(), off);
	d.padding = g.renderer().meshIndexOf(g, off);
	return d;
}

Drawable::Drawable(Geometry& g, uint32_t off, MatrixList& ml, StateSet& ss) : _matrixList(&ml)
{
	ss.appendDrawableInternal(*this, gpuDataOf(g, off, ml, nullptr));
}

Drawable::Drawable(Geometry& g, uint32_t off, MatrixList& ml, DataAllocation& dd, StateSet& ss)
    : _matrixList(&ml), _drawableData(&dd)
{
	ss.appendDrawableInternal(*this, gpuDataOf(g, off, ml, &dd));
}

Drawable::Drawable(Drawable&& o) noexcept
    : _stateSet(o._stateSet), _matrixList(o._matrixList), _drawableData(o._drawableData),
      _indexIntoStateSet(o._indexIntoStateSet)
{
	if (_indexIntoStateSet != ~0u) _stateSet->_drawablePtrList[_indexIntoStateSet] = this;
	o._indexIntoStateSet = ~0u;
}

Drawable& Drawable::operator=(Drawable&& rhs) noexcept
{
	if (this == &rhs) return *this;
	if (_indexIntoStateSet != ~0u) _stateSet->removeDrawableInternal(*this);
	_stateSet = rhs._stateSet;
	_matrixList = rhs._matrixList;
	_drawableData = rhs._drawableData;
	_indexIntoStateSet = rhs._indexIntoStateSet;
	if (_indexIntoStateSet != ~0u) _stateSet->_drawablePtrList[_indexIntoStateSet] = this;
	rhs._indexIntoStateSet = ~0u;
	return *this;
}

Drawable::~Drawable() noexcept
{
	if (_indexIntoStateSet != ~0u) _stateSet->removeDrawableInternal(*this);
}

void Drawable::destroy() noexcept
{
	if (_indexIntoStateSet != ~0u) {
		_stateSet->removeDrawableInternal(*this);
		_indexIntoStateSet = ~0u;
	}
}

// Drawable::create (Drawable.cpp:108-150): same StateSet -> only the record
// is rewritten in place; another StateSet -> removed there, appended here.
// As in the reference, the in-place update does not carry drawableData.
void Drawable::createInternal(Geometry& g, uint32_t off, MatrixList& ml, DataAllocation* dd, StateSet& ss)
{
	_matrixList = &ml;
	_drawableData = dd;
	if (_indexIntoStateSet != ~0u) {
		if (_stateSet == &ss) {
			_stateSet->_drawableDataList[_indexIntoStateSet] = gpuDataOf(g, off, ml, nullptr);
			_stateSet->markDirty(_indexIntoStateSet);
			return;
		}
		_stateSet->removeDrawableInternal(*this);
	}
	ss.appendDrawableInternal(*this, gpuDataOf(g, off, ml, nullptr));
}

void Drawable::create(Geometry& g, uint32_t off, MatrixList& ml, StateSet& ss) { createInternal(g, off, ml, nullptr, ss); }

void Drawable::create(Geometry& g, uint32_t off, MatrixList& ml, DataAllocation& dd, StateSet& ss)
{
	createInternal(g, off, ml, &dd, ss);
}

// -------------------------------------------------------------- StateSet

// Edges of the parent-child relation: one per link, in the parent's
// childList and the child's parentList (ParentChildList.h).
struct StateSetEdge {
	StateSet* parent;
	StateSet* child;
	std::list<StateSetEdge*>::iterator inParent, inChild;
};

template <bool CHILDREN>
StateSet& StateSetList<CHILDREN>::iterator::operator*() const
{
	return CHILDREN ? *(*_it)->child : *(*_it)->parent;
}

template <bool CHILDREN>
typename StateSetList<CHILDREN>::iterator StateSetList<CHILDREN>::append(StateSet& other)
{
	auto* e = new StateSetEdge{CHILDREN ? _owner : &other, CHILDREN ? &other : _owner, {}, {}};
	auto& children = e->parent->childList._list;
	auto& parents = e->child->parentList._list;
	try {
		children.push_back(e);
	} catch (...) {
		delete e;
		throw;
	}
	e->inParent = std::prev(children.end());
	try {
		parents.push_back(e);
	} catch (...) {
		children.erase(e->inParent);
		delete e;
		throw;
	}
	e->inChild = std::prev(parents.end());
	return iterator(CHILDREN ? e->inParent : e->inChild);
}

template <bool CHILDREN>
void StateSetList<CHILDREN>::remove(iterator it) noexcept
{
	StateSetEdge* e = *it.raw();
	e->parent->childList._list.erase(e->inParent);
	e->child->parentList._list.erase(e->inChild);
	delete e;
}

template class StateSetList<true>;
template class StateSetList<false>;

// The lists remove this StateSet's edges when they are destroyed.
StateSet::~StateSet() noexcept { removeAllDrawables(); }

void StateSet::markDirty(uint32_t index) noexcept
{
	_renderer->sceneChanged();
	if (_allDirty) return;
	if (_dirty.size() >= std::max<size_t>(64, _drawableDataList.size() / 4)) {
		_allDirty = true;
		_dirty.clear();
		return;
	}
	try {
		_dirty.push_back(index);
	} catch (...) {
		_allDirty = true;
		_dirty.clear();
	}
}

void StateSet::appendDrawableInternal(Drawable& d, const DrawableGpuData& gpuData)
{
	d._stateSet = this;
	d._indexIntoStateSet = uint32_t(_drawableDataList.size());
	_drawableDataList.emplace_back(gpuData);
	_drawablePtrList.emplace_back(&d);
	markDirty(d._indexIntoStateSet);
}

// Removal moves the last drawable into the hole (StateSet.cpp:29-47), so the
// order inside a StateSet is vector order, not creation order.
void StateSet::removeDrawableInternal(Drawable& d) noexcept
{
	const uint32_t i = d._indexIntoStateSet;
	const size_t last = _drawableDataList.size() - 1;
	if (i != last) {
		_drawableDataList[i] = _drawableDataList[last];
		Drawable* moved = _drawablePtrList[last];
		_drawablePtrList[i] = moved;
		moved->_indexIntoStateSet = i;
		markDirty(i);
	}
	_drawableDataList.pop_back();
	_drawablePtrList.pop_back();
	_renderer->sceneChanged();
}

void StateSet::appendDrawable(Drawable& d, const DrawableGpuData& gpuData)
{
	if (d._indexIntoStateSet != ~0u) d._stateSet->removeDrawableInternal(d);
	appendDrawableInternal(d, gpuData);
}

void StateSet::removeDrawable(Drawable& d)
{
	if (d._indexIntoStateSet == ~0u) return;
	d._stateSet->removeDrawableInternal(d);
	d._indexIntoStateSet = ~0u;
}

void StateSet::removeAllDrawables() noexcept
{
	for (Drawable* d : _drawablePtrList) d->_indexIntoStateSet = ~0u;
	_drawableDataList.clear();
	_drawablePtrList.clear();
	_renderer->sceneChanged();
}

// StateSet::prepareRecording (StateSet.cpp:180-195): drawables of this
// StateSet and its subtree, children visited once per parent edge.
size_t StateSet::prepareRecording()
{
	_skipRecording = !_forceRecording;
	size_t n = _drawableDataList.size();
	for (StateSet& c : childList) {
		n += c.prepareRecording();
		_skipRecording = _skipRecording && c._skipRecording;
	}
	_skipRecording = _skipRecording && n == 0;
	return n;
}

// StateSet::recordToCommandBuffer (StateSet.cpp:198-268): this StateSet's
// drawables go to [drawableCounter, drawableCounter + n), then the children.
void StateSet::recordToCommandBuffer(DrawableGpuData* staging, std::vector<DrawRange>& ranges, size_t& counter)
{
	if (_skipRecording) return;
	const size_t n = _drawableDataList.size();
	if (n > 0) {
		if (staging) std::memcpy(staging + counter, _drawableDataList.data(), n * sizeof(DrawableGpuData));
		ranges.push_back(DrawRange{this, counter, n});
		counter += n;
	}
	for (StateSet& c : childList) c.recordToCommandBuffer(staging, ranges, counter);
}

}  // namespace CadR
