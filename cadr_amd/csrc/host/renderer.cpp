// Renderer of the CadR facade: frame API, per-frame buffers, the gather
// dispatch and frame timing.
#include "../../../include/CadR/Renderer.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

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
	_dataStorage.reset();
	if (_defaultRenderer == this) _defaultRenderer = nullptr;
}

void Renderer::releaseBuffers() noexcept
{
	if (_device >= 0) {
		(void)hipFree(reinterpret_cast<void*>(_drawableBuffer));
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
	_drawableBuffer = _indirectBuffer = _pointersBuffer = 0;
	_staging = nullptr;
	_capacity = 0;
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
	// staging list in StateSet DFS order (StateSet.cpp:233-237)
	_pendingRanges.clear();
	size_t counter = 0;
	if (_collectFrameInfo) _inProgress.cpuRecordStateSetsBegin = getCpuTimestamp();
	if (_preparedRoot) _preparedRoot->recordToCommandBuffer(_staging, _pendingRanges, counter);
	if (_collectFrameInfo) _inProgress.cpuRecordStateSetsEnd = getCpuTimestamp();
	if (counter != numDrawables)
		throw LogicError("CadR::Renderer::recordDrawableProcessing(): numDrawables differs from the prepared scene");

	if (numDrawables == 0) {
		// two timestamps stand in for the skipped work (Renderer.cpp:600-620)
		recordEvent(1, stream);
		recordEvent(2, stream);
		return;
	}
	if (_device < 0) throw LogicError("CadR::Renderer::recordDrawableProcessing(): the renderer has no device");
	hipCheck(hipMemcpyAsync(reinterpret_cast<void*>(_drawableBuffer), _staging, numDrawables * sizeof(DrawableGpuData),
	                        hipMemcpyHostToDevice, hipStream_t(stream)),
	         "hipMemcpyAsync(drawable list)");
	recordEvent(1, stream);
	const cadrProcessDrawablesPushConstants pc{_dataStorage->handleTableDeviceAddress(), _drawableBuffer,
	                                           _indirectBuffer, _pointersBuffer};
	throwOnError(cadrProcessDrawables(&pc, _dataStorage->handleLevel(), numDrawables, stream),
	             "CadR::Renderer::recordDrawableProcessing");
	recordEvent(2, stream);
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
