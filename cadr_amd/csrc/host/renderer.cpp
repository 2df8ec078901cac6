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
	recordEvent(1, stream);
	const cadrProcessDrawablesPushConstants pc{_dataStorage->handleTableDeviceAddress(), _drawableBuffer,
	                                           _indirectBuffer, _pointersBuffer};
	throwOnError(cadrProcessDrawables(&pc, _dataStorage->handleLevel(), numDrawables, stream),
	             "CadR::Renderer::recordDrawableProcessing");
	recordEvent(2, stream);
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
