// Renderer of the CadR facade: frame API, per-frame buffers, the gather
// dispatch and frame timing.
#include "../../../include/CadR/Renderer.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>

#include "../../../include/CadR/Exceptions.h"
#include "../../../include/cadr_hip.h"

namespace CadR {

Renderer* Renderer::_defaultRenderer = nullptr;

static void hipCheck(hipError_t e, const char* what)
{
	if (e != hipSuccess) {
		(void)hipGetLastError();
		if (e == hipErrorOutOfMemory) throw OutOfResources(std::string(what) + ": " + hipGetErrorString(e));
		throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
	}
}

Renderer::Renderer(bool makeDefault) : _dataStorage(new DataStorage(*this, hostMemoryProvider()))
{
	if (makeDefault) _defaultRenderer = this;
}

Renderer::Renderer(int deviceIndex, bool makeDefault) : _device(deviceIndex)
{
	hipCheck(hipSetDevice(deviceIndex), "hipSetDevice");
	_dataStorage.reset(new DataStorage(*this, hipMemoryProvider()));
	for (auto& e : _events) {
		hipEvent_t ev = nullptr;
		hipCheck(hipEventCreate(&ev), "hipEventCreate");
		e = ev;
	}
	for (void** e : {&_uploadDone, &_listFree}) {
		hipEvent_t ev = nullptr;
		hipCheck(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
		*e = ev;
	}
	if (makeDefault) _defaultRenderer = this;
}

Renderer::~Renderer() { finalize(); }

void Renderer::finalize()
{
	if (_device >= 0) (void)hipDeviceSynchronize();
	releaseBuffers();
	for (auto& e : _events) {
		if (e) (void)hipEventDestroy(hipEvent_t(e));
		e = nullptr;
	}
	for (void** e : {&_uploadDone, &_listFree}) {
		if (*e) (void)hipEventDestroy(hipEvent_t(*e));
		*e = nullptr;
	}
	_uploadPending = _listFreePending = false;
	releaseCull();
	if (_delta) (void)hipHostFree(_delta);
	if (_deltaDevice) (void)hipFree(reinterpret_cast<void*>(_deltaDevice));
	_delta = nullptr;
	_deltaDevice = 0;
	_deltaCapacity = 0;
	_dataStorage.reset();
	if (_defaultRenderer == this) _defaultRenderer = nullptr;
}

void Renderer::releaseBuffers() noexcept
{
	if (_device >= 0) {
		(void)hipFree(reinterpret_cast<void*>(_drawableBuffer));
		(void)hipFree(reinterpret_cast<void*>(_drawableBufferSpare));
		(void)hipFree(reinterpret_cast<void*>(_indirectBuffer));
		(void)hipFree(reinterpret_cast<void*>(_pointersBuffer));
		if (_staging) (void)hipHostFree(_staging);
	} else {
		MemoryProvider& p = hostMemoryProvider();
		if (_drawableBuffer) p.release(_drawableBuffer);
		if (_indirectBuffer) p.release(_indirectBuffer);
		if (_pointersBuffer) p.release(_pointersBuffer);
		std::free(_staging);
	}
	_drawableBuffer = _drawableBufferSpare = _indirectBuffer = _pointersBuffer = 0;
	_staging = nullptr;
	_capacity = 0;
	_residentValid = false;
	_residentRanges.clear();
}

void Renderer::setResidentDrawables(bool on)
{
	_resident = on;
	_residentValid = false;
}

uint64_t Renderer::getCpuTimestamp()
{
	return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
	                    std::chrono::steady_clock::now().time_since_epoch())
	                    .count());
}

size_t Renderer::beginFrame()
{
	_frameNumber++;
	if (_collectFrameInfo) {
		_inProgress = FrameInfo{};
		_inProgress.frameNumber = _frameNumber;
		_inProgress.cpuBeginFrame = getCpuTimestamp();
		for (bool& r : _eventRecorded) r = false;
	}
	return _frameNumber;
}

// Uploads staged data and handle tables, synchronously like the reference's
// submit-and-wait (Renderer.cpp:946-999).
void Renderer::executeCopyOperations()
{
	_dataStorage->recordUploads();
}

void Renderer::recordEvent(int index, void* stream)
{
	if (!_collectFrameInfo || _device < 0) return;
	hipCheck(hipEventRecord(hipEvent_t(_events[index]), hipStream_t(stream)), "hipEventRecord");
	_eventRecorded[index] = true;
}

void Renderer::beginRecording(void* stream) { recordEvent(0, stream); }

size_t Renderer::prepareSceneRendering(StateSet& root)
{
	if (_collectFrameInfo) _inProgress.cpuPrepareRecordingBegin = getCpuTimestamp();
	const size_t n = root.prepareRecording();
	if (_collectFrameInfo) _inProgress.cpuPrepareRecordingEnd = getCpuTimestamp();
	_preparedRoot = &root;

	// grow the buffers to 1.2 N, at least 128 slots (Renderer.cpp:461-470)
	if (_capacity < n) {
		size_t slots = size_t(float(n) * 1.2f);
		if (slots < 128) slots = 128;
		if (_device >= 0) {
			// the last frame may still copy out of the staging list or read the buffers
			if (_uploadPending) hipCheck(hipEventSynchronize(hipEvent_t(_uploadDone)), "hipEventSynchronize");
			if (_listFreePending) hipCheck(hipEventSynchronize(hipEvent_t(_listFree)), "hipEventSynchronize");
			_uploadPending = _listFreePending = false;
		}
		releaseBuffers();
		if (_device >= 0) {
			void *d = nullptr, *i = nullptr, *p = nullptr;
			hipCheck(hipMalloc(&d, slots * sizeof(DrawableGpuData)), "hipMalloc(drawable buffer)");
			hipCheck(hipMalloc(&i, slots * sizeof(cadrDrawIndirectCommand)), "hipMalloc(indirect buffer)");
			hipCheck(hipMalloc(&p, slots * drawablePointersRecordSize), "hipMalloc(drawable pointers buffer)");
			hipCheck(hipHostMalloc(reinterpret_cast<void**>(&_staging), slots * sizeof(DrawableGpuData),
			                       hipHostMallocDefault),
			         "hipHostMalloc(drawable staging)");
			_drawableBuffer = reinterpret_cast<uint64_t>(d);
			_indirectBuffer = reinterpret_cast<uint64_t>(i);
			_pointersBuffer = reinterpret_cast<uint64_t>(p);
		} else {
			MemoryProvider& mp = hostMemoryProvider();
			_drawableBuffer = mp.allocate(slots * sizeof(DrawableGpuData));
			_indirectBuffer = mp.allocate(slots * sizeof(cadrDrawIndirectCommand));
			_pointersBuffer = mp.allocate(slots * drawablePointersRecordSize);
			_staging = static_cast<DrawableGpuData*>(std::malloc(slots * sizeof(DrawableGpuData)));
			if (!_staging) throw OutOfResources("CadR::Renderer: cannot allocate the drawable staging list");
		}
		_capacity = slots;
	}
	return n;
}

void Renderer::recordDrawableProcessing(void* stream, size_t numDrawables)
{
	if (numDrawables > _capacity)
		throw LogicError("CadR::Renderer::recordDrawableProcessing(): more drawables than prepareSceneRendering() counted");
	const hipStream_t s = hipStream_t(stream);
	// the host staging and delta lists are rewritten below: the last frame's
	// copies out of them must have run
	if (_uploadPending) {
		hipCheck(hipEventSynchronize(hipEvent_t(_uploadDone)), "hipEventSynchronize");
		_uploadPending = false;
	}
	// staging list in StateSet DFS order (StateSet.cpp:233-237); a resident
	// list needs only the ranges
	const bool resident = _resident && _device >= 0 && _residentValid;
	_pendingRanges.clear();
	size_t counter = 0;
	if (_collectFrameInfo) _inProgress.cpuRecordStateSetsBegin = getCpuTimestamp();
	if (_preparedRoot) _preparedRoot->recordToCommandBuffer(resident ? nullptr : _staging, _pendingRanges, counter);
	if (_collectFrameInfo) _inProgress.cpuRecordStateSetsEnd = getCpuTimestamp();
	if (counter != numDrawables)
		throw LogicError("CadR::Renderer::recordDrawableProcessing(): numDrawables differs from the prepared scene");

	_lastUploadBytes = 0;
	_processedFrame = _frameNumber;
	_processedCount = numDrawables;
	if (numDrawables == 0) {
		// two timestamps stand in for the skipped work (Renderer.cpp:600-620)
		recordEvent(1, stream);
		recordEvent(2, stream);
		return;
	}
	if (_device < 0) throw LogicError("CadR::Renderer::recordDrawableProcessing(): the renderer has no device");
	// the last frame's gather (perhaps on another stream) reads the list
	if (_listFreePending) hipCheck(hipStreamWaitEvent(s, hipEvent_t(_listFree), 0), "hipStreamWaitEvent");
	if (resident)
		uploadResidentList(stream, _pendingRanges);
	else
		uploadFullList(stream, numDrawables);
	hipCheck(hipEventRecord(hipEvent_t(_uploadDone), s), "hipEventRecord");
	_uploadPending = true;
	// the timestamps around the dispatch (Renderer.cpp:660-700) are recorded
	// by the dispatch itself: no marker commands between the upload and the
	// gather
	if (_collectFrameInfo)
		throwOnError(cadrTimeNextLaunch(_events[1], _events[2]), "CadR::Renderer::recordDrawableProcessing");
	const cadrProcessDrawablesPushConstants pc{_dataStorage->handleTableDeviceAddress(), _drawableBuffer,
	                                           _indirectBuffer, _pointersBuffer};
	const int rc = cadrProcessDrawables(&pc, _dataStorage->handleLevel(), numDrawables, stream);
	if (rc != CADR_SUCCESS) {
		if (_collectFrameInfo) (void)cadrTimeNextLaunch(nullptr, nullptr);
		throwOnError(rc, "CadR::Renderer::recordDrawableProcessing");
	}
	if (_collectFrameInfo) _eventRecorded[1] = _eventRecorded[2] = true;
	hipCheck(hipEventRecord(hipEvent_t(_listFree), s), "hipEventRecord");
	_listFreePending = true;
}

// The whole staged list to the device (Renderer.cpp:635-644); in resident
// mode it becomes the resident list.
void Renderer::uploadFullList(void* stream, size_t numDrawables)
{
	hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(_drawableBuffer), _staging, numDrawables * sizeof(DrawableGpuData),
	                        hipMemcpyHostToDevice, hipStream_t(stream)),
	         "hipMemcpyAsync(drawable list)");
	_lastUploadBytes = numDrawables * sizeof(DrawableGpuData);
	if (_resident) {
		_residentRanges = _pendingRanges;
		_residentValid = true;
		for (const DrawRange& r : _pendingRanges) {
			r.stateSet->_dirty.clear();
			r.stateSet->_allDirty = false;
		}
	}
}

void Renderer::reserveDelta(size_t count)
{
	if (count <= _deltaCapacity) return;
	const size_t cap = std::max<size_t>({count, 2 * _deltaCapacity, 1024});
	const size_t bytes = cap * (sizeof(DrawableGpuData) + sizeof(uint32_t));
	if (_delta) (void)hipHostFree(_delta);
	if (_deltaDevice) (void)hipFree(reinterpret_cast<void*>(_deltaDevice));
	_delta = nullptr;
	_deltaDevice = 0;
	_deltaCapacity = 0;
	void *h = nullptr, *d = nullptr;
	hipCheck(hipHostMalloc(&h, bytes, hipHostMallocDefault), "hipHostMalloc(drawable deltas)");
	_delta = static_cast<uint8_t*>(h);
	hipCheck(hipMalloc(&d, bytes), "hipMalloc(drawable deltas)");
	_deltaDevice = reinterpret_cast<uint64_t>(d);
	_deltaCapacity = cap;
}

// Brings the resident list to this frame's ranges: changed records as
// deltas; when the ranges moved, a rebuild into the spare list from the
// StateSets the resident list holds; StateSets new to the list (or with too
// many changes to list) whole through the staging list.
void Renderer::uploadResidentList(void* stream, const std::vector<DrawRange>& ranges)
{
	const hipStream_t s = hipStream_t(stream);
	constexpr size_t R = sizeof(DrawableGpuData);
	bool same = ranges.size() == _residentRanges.size();
	for (size_t i = 0; same && i < ranges.size(); i++)
		same = ranges[i].stateSet == _residentRanges[i].stateSet && ranges[i].base == _residentRanges[i].base &&
		       ranges[i].count == _residentRanges[i].count;
	std::unordered_map<const StateSet*, DrawRange> held;
	uint64_t dst = _drawableBuffer;
	if (!same) {
		if (!_drawableBufferSpare) {
			void* p = nullptr;
			hipCheck(hipMalloc(&p, _capacity * R), "hipMalloc(drawable buffer)");
			_drawableBufferSpare = reinterpret_cast<uint64_t>(p);
		}
		dst = _drawableBufferSpare;
		for (const DrawRange& r : _residentRanges) held.emplace(r.stateSet, r);
	}
	auto whole = [&](const DrawRange& r) { return r.stateSet->_allDirty || (!same && !held.count(r.stateSet)); };

	size_t deltas = 0;
	for (const DrawRange& r : ranges)
		if (!whole(r))
			for (uint32_t i : r.stateSet->_dirty) deltas += i < r.count;
	reserveDelta(deltas);
	auto* records = reinterpret_cast<DrawableGpuData*>(_delta);
	auto* slots = reinterpret_cast<uint32_t*>(_delta + deltas * R);

	size_t k = 0, sent = 0;
	for (const DrawRange& r : ranges) {
		const StateSet& ss = *r.stateSet;
		if (whole(r)) {
			std::memcpy(_staging + r.base, ss._drawableDataList.data(), r.count * R);
			hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(dst + r.base * R), _staging + r.base, r.count * R,
			                        hipMemcpyHostToDevice, s),
			         "hipMemcpyAsync(drawable list)");
			sent += r.count * R;
			continue;
		}
		if (!same) {
			const DrawRange& old = held.at(r.stateSet);
			const size_t keep = std::min(old.count, r.count);
			if (keep)
				hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(dst + r.base * R),
				                        reinterpret_cast<const void*>(_drawableBuffer + old.base * R), keep * R,
				                        hipMemcpyDeviceToDevice, s),
				         "hipMemcpyAsync(drawable list move)");
		}
		for (uint32_t i : ss._dirty) {
			if (i >= r.count) continue;
			records[k] = ss._drawableDataList[i];
			slots[k] = uint32_t(r.base + i);
			k++;
		}
	}
	for (const DrawRange& r : ranges) {
		r.stateSet->_dirty.clear();
		r.stateSet->_allDirty = false;
	}
	if (k) {
		const size_t bytes = k * (R + sizeof(uint32_t));
		hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(_deltaDevice), _delta, bytes, hipMemcpyHostToDevice, s),
		         "hipMemcpyAsync(drawable deltas)");
		throwOnError(cadrScatterDrawables(dst, _capacity, _deltaDevice, _deltaDevice + k * R, uint32_t(k), stream),
		             "CadR::Renderer::recordDrawableProcessing");
		sent += bytes;
	}
	if (!same) std::swap(_drawableBuffer, _drawableBufferSpare);
	_residentRanges = ranges;
	_lastUploadBytes = sent;
}

void Renderer::recordSceneRendering(void*, StateSet& root, std::vector<DrawRange>& ranges)
{
	if (&root != _preparedRoot) throw LogicError("CadR::Renderer::recordSceneRendering(): root was not prepared");
	ranges = _pendingRanges;
	if (_culledFrame != _frameNumber) {
		for (DrawRange& r : ranges) {  // no cull frame: the items are the drawables
			r.itemBase = r.base;
			r.itemCount = r.count;
		}
	}
}

// ------------------------------------------------------------ cull frame

uint32_t Renderer::meshIndexOf(const Geometry& g, uint32_t primitiveSetOffset)
{
	const auto key = std::make_pair(&g, primitiveSetOffset);
	const auto it = _meshIndex.find(key);
	if (it != _meshIndex.end()) return it->second;
	const uint32_t m = uint32_t(_meshKeys.size());
	_meshKeys.push_back(key);
	_meshIndex.emplace(key, m);
	return m;
}

void Renderer::forgetGeometry(const Geometry* g) noexcept
{
	auto it = _meshIndex.lower_bound(std::make_pair(g, 0u));
	while (it != _meshIndex.end() && it->first.first == g) {
		_meshKeys[it->second].first = nullptr;
		it = _meshIndex.erase(it);
	}
}

void Renderer::setCullConfig(const cadrCullConfig& singleView, const cadrCullConfig& batched)
{
	if (!(cadrCullConfigPaths(&singleView) & CADR_CULL_PATH_VIEW) ||
	    !(cadrCullConfigPaths(&batched) & CADR_CULL_PATH_BATCH))
		throw LogicError("CadR::Renderer::setCullConfig(): unsupported blockSize / itemsPerThread");
	_cullSingle = singleView;
	_cullBatched = batched;
}

void Renderer::growCull(uint64_t& buffer, size_t& capacity, size_t bytes, const char* what)
{
	if (bytes <= capacity && buffer) return;
	(void)hipFree(reinterpret_cast<void*>(buffer));
	buffer = 0;
	capacity = 0;
	const size_t cap = std::max<size_t>(256, bytes + bytes / 5);
	void* p = nullptr;
	hipCheck(hipMalloc(&p, cap), what);
	buffer = reinterpret_cast<uint64_t>(p);
	capacity = cap;
}

// The same growth rule for a buffer that a Vulkan renderer imports: a
// shared buffer (cadrSharedBufferCreate), replaced once the device is idle,
// the generation telling the renderer to import it again.
void Renderer::growCullShared(uint64_t& buffer, size_t& capacity, cadrSharedBuffer*& shared, size_t bytes,
                              const char* what)
{
	if (bytes <= capacity && buffer) return;
	if (shared) {
		hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
		cadrSharedBufferDestroy(shared);
	} else if (buffer) {
		(void)hipFree(reinterpret_cast<void*>(buffer));
	}
	shared = nullptr;
	buffer = 0;
	capacity = 0;
	const size_t cap = std::max<size_t>(256, bytes + bytes / 5);
	throwOnError(cadrSharedBufferCreate(cap, &shared), what);
	buffer = cadrSharedBufferAddress(shared);
	capacity = size_t(cadrSharedBufferSize(shared));
	_cullGeneration++;
}

void Renderer::setCullBuffersShared(bool on)
{
	if (on == _cullShared) return;
	if (_device < 0 && on)
		throw LogicError("CadR::Renderer::setCullBuffersShared(): a device-less renderer shares nothing");
	// the next cull frame allocates the output buffers of the new kind
	if (_device >= 0) hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
	for (auto bc : {std::make_pair(&_cull.commands, &_cull.sharedCommands),
	                std::make_pair(&_cull.counts, &_cull.sharedCounts),
	                std::make_pair(&_cull.pointers, &_cull.sharedPointers)}) {
		if (*bc.second) cadrSharedBufferDestroy(*bc.second);
		else if (*bc.first) (void)hipFree(reinterpret_cast<void*>(*bc.first));
		*bc.first = 0;
		*bc.second = nullptr;
	}
	_cull.commandsCap = _cull.countsCap = _cull.pointersCap = 0;
	_cullShared = on;
	_cullGeneration++;
}

int Renderer::exportCullFd(const cadrSharedBuffer* shared, const char* what) const
{
	if (!_cullShared || !shared)
		throw LogicError(std::string("CadR::Renderer::") + what +
		                 "(): no shared buffer (setCullBuffersShared, then a cull frame)");
	int fd = -1;
	throwOnError(cadrSharedBufferExportFd(shared, &fd), what);
	return fd;
}

int Renderer::exportCullCommandBufferFd() const { return exportCullFd(_cull.sharedCommands, "exportCullCommandBufferFd"); }
int Renderer::exportCullCountBufferFd() const { return exportCullFd(_cull.sharedCounts, "exportCullCountBufferFd"); }
int Renderer::exportCullPointersBufferFd() const
{
	return exportCullFd(_cullPointersOn ? _cull.sharedPointers : nullptr, "exportCullPointersBufferFd");
}

void Renderer::releaseCull() noexcept
{
	// shared output buffers are released as such, the rest with hipFree
	for (auto bc : {std::make_pair(&_cull.commands, &_cull.sharedCommands),
	                std::make_pair(&_cull.counts, &_cull.sharedCounts),
	                std::make_pair(&_cull.pointers, &_cull.sharedPointers)}) {
		if (!*bc.second) continue;
		cadrSharedBufferDestroy(*bc.second);
		*bc.second = nullptr;
		*bc.first = 0;
	}
	for (uint64_t* b : {&_cull.spheres, &_cull.meshIndex, &_cull.itemBase, &_cull.itemSource, &_cull.rangeBase,
	                    &_cull.tiles, &_cull.upload, &_cull.commands, &_cull.counts, &_cull.pointers, &_cull.meshes})
		if (*b) (void)hipFree(reinterpret_cast<void*>(*b));
	_cull = CullBuffers{};
	if (_cullCopied) (void)hipEventDestroy(hipEvent_t(_cullCopied));
	_cullCopied = nullptr;
	if (_cullPinned) (void)hipHostFree(_cullPinned);
	_cullPinned = nullptr;
	_cullPinnedCap = 0;
	_cullCopyPending = false;
	if (_cullWs) cadrCullWorkspaceDestroy(_cullWs);
	_cullWs = nullptr;
	_cullWsWords = 0;
	_cullLayoutEpoch = 0;
	_cullRanges.clear();
	_cullTileSize = 0;
}

// Item layout of the frame's ranges: one cull item per matrix of each
// drawable's MatrixList, in slot order.  Rebuilt only when a drawable or a
// MatrixList size changed (scene epoch) or the ranges moved; with exactly one
// matrix per drawable the items are the drawable slots themselves.
void Renderer::updateCullLayout(void* stream)
{
	(void)stream;
	bool same = _cullLayoutEpoch == _sceneEpoch && _cullRanges.size() == _pendingRanges.size();
	for (size_t k = 0; same && k < _pendingRanges.size(); k++)
		same = _cullRanges[k].stateSet == _pendingRanges[k].stateSet && _cullRanges[k].base == _pendingRanges[k].base &&
		       _cullRanges[k].count == _pendingRanges[k].count;
	if (same) {
		_pendingRanges = _cullRanges;
		return;
	}
	bool identity = true;
	for (const DrawRange& r : _pendingRanges)
		for (size_t j = 0; identity && j < r.count; j++) identity = r.stateSet->getDrawable(j).matrixList()->numMatrices() == 1;
	size_t items = 0;
	std::vector<uint32_t> itemBase, itemDrawable, rangeBase(std::max<size_t>(_pendingRanges.size(), 1), 0u);
	if (!identity) itemBase.resize(_processedCount);
	for (size_t k = 0; k < _pendingRanges.size(); k++) {
		DrawRange& r = _pendingRanges[k];
		r.itemBase = items;
		if (identity) {
			items += r.count;
		} else {
			for (size_t j = 0; j < r.count; j++) {
				itemBase[r.base + j] = uint32_t(items);
				const size_t m = r.stateSet->getDrawable(j).matrixList()->numMatrices();
				itemDrawable.insert(itemDrawable.end(), m, uint32_t(r.base + j));
				items += m;
			}
		}
		r.itemCount = items - r.itemBase;
		rangeBase[k] = uint32_t(r.itemBase);
		if (items >= 0xffffffffull) throw OutOfResources("CadR::Renderer::recordDrawableCulling(): more than 2^32 - 1 cull items");
	}
	if (!identity) {
		growCull(_cull.itemBase, _cull.itemBaseCap, itemBase.size() * 4, "hipMalloc(cull item bases)");
		hipCheck(hipMemcpy(reinterpret_cast<void*>(_cull.itemBase), itemBase.data(), itemBase.size() * 4,
		                   hipMemcpyHostToDevice),
		         "hipMemcpy(cull item bases)");
		growCull(_cull.itemSource, _cull.itemSourceCap, std::max<size_t>(items, 1) * 4, "hipMalloc(cull item drawables)");
		if (items)
			hipCheck(hipMemcpy(reinterpret_cast<void*>(_cull.itemSource), itemDrawable.data(), items * 4,
			                   hipMemcpyHostToDevice),
			         "hipMemcpy(cull item drawables)");
	}
	growCull(_cull.rangeBase, _cull.rangeBaseCap, rangeBase.size() * 4, "hipMalloc(cull range bases)");
	hipCheck(hipMemcpy(reinterpret_cast<void*>(_cull.rangeBase), rangeBase.data(), rangeBase.size() * 4,
	                   hipMemcpyHostToDevice),
	         "hipMemcpy(cull range bases)");
	growCull(_cull.spheres, _cull.spheresCap, items * 16, "hipMalloc(cull spheres)");
	growCull(_cull.meshIndex, _cull.meshIndexCap, items * 4, "hipMalloc(cull mesh indices)");
	_cullIdentity = identity;
	_cullItems = items;
	_cullRanges = _pendingRanges;
	_cullLayoutEpoch = _sceneEpoch;
	_cullTileSize = 0;  // tiles follow the item counts
}

void Renderer::recordDrawableCulling(void* stream, const cadrView* views, size_t numViews)
{
	const char* fn = "CadR::Renderer::recordDrawableCulling";
	if (_device < 0) throw LogicError(std::string(fn) + "(): the renderer has no device");
	if (_processedFrame != _frameNumber)
		throw LogicError(std::string(fn) + "(): recordDrawableProcessing() of this frame must come first");
	if (numViews == 0 || !views) throw LogicError(std::string(fn) + "(): no views");
	if (numViews > 0xffffffffull) throw LogicError(std::string(fn) + "(): too many views");
	const hipStream_t s = hipStream_t(stream);
	const cadrCullConfig& cfg = numViews == 1 ? _cullSingle : _cullBatched;
	const uint32_t tile = cadrCullTileSize(&cfg);
	updateCullLayout(stream);
	const size_t N = _processedCount, items = _cullItems, K = _cullRanges.size(), V = numViews;
	_cullViews = V;
	_culledFrame = _frameNumber;
	if (K == 0) return;
	// the host copies below are read by the last frame's async copies until they ran
	if (!_cullCopied) {
		hipEvent_t ev = nullptr;
		hipCheck(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
		_cullCopied = ev;
	}
	if (_cullCopyPending) hipCheck(hipEventSynchronize(hipEvent_t(_cullCopied)), "hipEventSynchronize");
	_cullCopyPending = false;
	// pinned staging: job, views, mesh sources, sent as one copy into one
	// device region (pageable copies would wait for the stream; three small
	// copies cost a blit launch each)
	const size_t M = _meshKeys.size();
	const size_t viewsAt = 128, sourcesAt = viewsAt + (V * sizeof(cadrView) + 127) / 128 * 128;
	const size_t pinnedBytes = sourcesAt + std::max<size_t>(M, 1) * sizeof(cadrMeshSource);
	if (_cullPinnedCap < pinnedBytes) {
		if (_cullPinned) (void)hipHostFree(_cullPinned);
		_cullPinned = nullptr;
		_cullPinnedCap = 0;
		void* p = nullptr;
		hipCheck(hipHostMalloc(&p, pinnedBytes + pinnedBytes / 4, hipHostMallocDefault), "hipHostMalloc(cull staging)");
		_cullPinned = static_cast<uint8_t*>(p);
		_cullPinnedCap = pinnedBytes + pinnedBytes / 4;
	}
	auto* src = reinterpret_cast<cadrMeshSource*>(_cullPinned + sourcesAt);
	auto* viewsHost = reinterpret_cast<cadrView*>(_cullPinned + viewsAt);
	cadrCullJob& job = *reinterpret_cast<cadrCullJob*>(_cullPinned);
	if (items * V > (size_t(1) << 40) / 20)
		throw OutOfResources(std::string(fn) + "(): the per-view command buffers would exceed 1 TiB");

	// mesh table: each (Geometry, PrimitiveSet offset) with its current sphere and LODs
	for (size_t m = 0; m < M; m++) {
		cadrMeshSource& o = src[m];
		o = cadrMeshSource{};
		const Geometry* g = _meshKeys[m].first;
		const uint32_t off = _meshKeys[m].second;
		if (!g) continue;  // gone: never visible
		const size_t psBytes = g->primitiveSetDataAllocation().size();
		const size_t fit = psBytes >= off + 8u ? (psBytes - off) / 8u : 0u;  // LOD records inside the allocation
		o.primitiveSetHandle = g->primitiveSetDataAllocation().handle();
		o.primitiveSetOffset = off;
		o.lodCount = uint32_t(std::min<size_t>(g->numLods(), fit));
		const BoundingSphere& b = g->boundingSphere();
		o.sphere[0] = b.center[0]; o.sphere[1] = b.center[1]; o.sphere[2] = b.center[2]; o.sphere[3] = b.radius;
	}
	std::memcpy(viewsHost, views, V * sizeof(cadrView));
	growCull(_cull.upload, _cull.uploadCap, pinnedBytes, "hipMalloc(cull job, views, mesh sources)");
	growCull(_cull.meshes, _cull.meshesCap, std::max<size_t>(M, 1) * sizeof(cadrMesh), "hipMalloc(mesh table)");

	// outputs, tiles and the workspace
	if (_cullShared) {
		growCullShared(_cull.commands, _cull.commandsCap, _cull.sharedCommands, std::max<size_t>(V * items * 20, 4),
		               "cadrSharedBufferCreate(cull commands)");
		growCullShared(_cull.counts, _cull.countsCap, _cull.sharedCounts, std::max<size_t>(V * K * 4, 4),
		               "cadrSharedBufferCreate(cull counts)");
		if (_cullPointersOn)
			growCullShared(_cull.pointers, _cull.pointersCap, _cull.sharedPointers,
			               std::max<size_t>(V * items * 32, 16), "cadrSharedBufferCreate(cull pointers)");
	} else {
		growCull(_cull.commands, _cull.commandsCap, std::max<size_t>(V * items * 20, 4), "hipMalloc(cull commands)");
		growCull(_cull.counts, _cull.countsCap, V * K * 4, "hipMalloc(cull counts)");
		if (_cullPointersOn)
			growCull(_cull.pointers, _cull.pointersCap, std::max<size_t>(V * items * 32, 16), "hipMalloc(cull pointers)");
	}
	if (_cullTileSize != tile) {
		std::vector<uint32_t> counts(K);
		for (size_t k = 0; k < K; k++) counts[k] = uint32_t(_cullRanges[k].itemCount);
		uint32_t nt = 0;
		throwOnError(cadrBuildCullTiles(counts.data(), uint32_t(K), tile, nullptr, 0, &nt), fn);
		std::vector<cadrCullTile> tiles(nt);
		throwOnError(cadrBuildCullTiles(counts.data(), uint32_t(K), tile, tiles.data(), nt, &nt), fn);
		growCull(_cull.tiles, _cull.tilesCap, nt * sizeof(cadrCullTile), "hipMalloc(cull tiles)");
		hipCheck(hipMemcpy(reinterpret_cast<void*>(_cull.tiles), tiles.data(), nt * sizeof(cadrCullTile),
		                   hipMemcpyHostToDevice),
		         "hipMemcpy(cull tiles)");
		_cullNumTiles = nt;
		_cullTileSize = tile;
	}
	const uint64_t words = uint64_t(V) * _cullNumTiles;
	if (!_cullWs || _cullWsWords < words) {
		if (_cullWs) {
			hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
			cadrCullWorkspaceDestroy(_cullWs);
			_cullWs = nullptr;
		}
		throwOnError(cadrCullWorkspaceCreate(words + words / 4, 1, &_cullWs), fn);
		_cullWsWords = words + words / 4;
	}
	job = cadrCullJob{};
	job.spheres = _cull.spheres;
	job.meshIndex = _cull.meshIndex;
	job.meshes = _cull.meshes;
	job.tiles = _cull.tiles;
	job.views = _cull.upload + viewsAt;
	job.commands = _cull.commands;
	job.counts = _cull.counts;
	job.numDrawables = uint32_t(items);
	job.numTiles = _cullNumTiles;
	job.numStateSets = uint32_t(K);
	job.numViews = uint32_t(V);
	job.commandStride = uint32_t(items);
	job.numMeshes = uint32_t(M);
	hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(_cull.upload), _cullPinned, pinnedBytes, hipMemcpyHostToDevice, s),
	         "hipMemcpyAsync(cull job, views, mesh sources)");
	hipCheck(hipEventRecord(hipEvent_t(_cullCopied), s), "hipEventRecord");
	_cullCopyPending = true;

	const uint64_t table = _dataStorage->handleTableDeviceAddress();
	const uint32_t level = _dataStorage->handleLevel();
	throwOnError(cadrBuildMeshTable(table, level, _cull.upload + sourcesAt, uint32_t(M), _cull.meshes, stream), fn);

	// cull items from the resident drawable list
	if (items > 0) {
		cadrCullItemsDesc d{};
		d.handleTable = table;
		d.drawableList = _drawableBuffer;
		d.itemBase = _cullIdentity ? 0 : _cull.itemBase;
		d.itemDrawable = _cullIdentity ? 0 : _cull.itemSource;
		d.meshes = _cull.meshes;
		d.spheres = _cull.spheres;
		d.meshIndex = _cull.meshIndex;
		d.numDrawables = N;
		d.numItems = items;
		d.handleLevel = level;
		d.numMeshes = uint32_t(M);
		throwOnError(cadrBuildCullItems(&d, stream), fn);
	}

	if (V == 1)
		throwOnError(cadrCullView(&cfg, _cull.upload, &job, _cullWs, stream), fn);
	else
		throwOnError(cadrCullBatch(&cfg, _cull.upload, &job, 1, _cullWs, stream), fn);

	// the reference's draw contract: instance index and DrawablePointers per command
	throwOnError(cadrFinishCullCommands(_cull.commands, _cull.counts, _cull.rangeBase, uint32_t(V), uint32_t(K),
	                                    uint32_t(items), _cullIdentity ? 0 : _cull.itemBase,
	                                    _cullIdentity ? 0 : _cull.itemSource, items, _pointersBuffer,
	                                    N, _cullPointersOn ? _cull.pointers : 0, stream),
	             fn);
	for (size_t k = 0; k < K; k++) {
		_pendingRanges[k].itemBase = _cullRanges[k].itemBase;
		_pendingRanges[k].itemCount = _cullRanges[k].itemCount;
	}
}

void Renderer::downloadCullCounts(std::vector<uint32_t>& counts) const
{
	counts.assign(_cullViews * _cullRanges.size(), 0u);
	if (counts.empty()) return;
	hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
	hipCheck(hipMemcpy(counts.data(), reinterpret_cast<const void*>(_cull.counts), counts.size() * 4,
	                   hipMemcpyDeviceToHost),
	         "hipMemcpy(cull counts)");
}

void Renderer::endRecording(void* stream) { recordEvent(3, stream); }

void Renderer::endFrame()
{
	if (_collectFrameInfo) {
		_inProgress.cpuEndFrame = getCpuTimestamp();
		_infoFrame = _frameNumber;
	}
}

const FrameInfo& Renderer::getFrameInfo()
{
	if (!_collectFrameInfo || _infoFrame != _frameNumber || _device < 0) return _completed;
	if (!_eventRecorded[0] || !_eventRecorded[3]) return _completed;
	hipCheck(hipEventSynchronize(hipEvent_t(_events[3])), "hipEventSynchronize");
	auto ns = [&](int i) -> uint64_t {
		if (!_eventRecorded[i]) return 0;
		float ms = 0.f;
		hipCheck(hipEventElapsedTime(&ms, hipEvent_t(_events[0]), hipEvent_t(_events[i])), "hipEventElapsedTime");
		return uint64_t(std::llround(double(ms) * 1e6));
	};
	_completed = _inProgress;
	_completed.gpuBeginExecution = 0;
	_completed.gpuAfterTransfersAndBeforeDrawableProcessing = ns(1);
	_completed.gpuAfterDrawableProcessingAndBeforeRendering = ns(2);
	_completed.gpuEndExecution = ns(3);
	return _completed;
}

}  // namespace CadR
