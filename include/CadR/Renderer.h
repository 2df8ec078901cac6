// Renderer of the CadR facade: the frame API around the drawable-processing
// pass, on one HIP device and caller-owned HIP streams.
//
// Mirrors CadR::Renderer (src/CadR/Renderer.h:34-201) for this path:
//   beginFrame -> executeCopyOperations -> beginRecording ->
//   prepareSceneRendering(root) -> recordDrawableProcessing(stream, n) ->
//   recordSceneRendering(stream, root, ranges) -> endRecording -> endFrame.
// prepareSceneRendering counts the drawables (StateSet::prepareRecording)
// and grows the four per-frame buffers to ceil(1.2 N), at least 128 slots
// (Renderer.cpp:446-595): drawable list 48 B, host staging 48 B, indirect
// commands 16 B, DrawablePointers 32 B.  recordDrawableProcessing fills the
// staging list in StateSet DFS order, copies it to the device and launches
// the gather kernel with the push constants of processDrawables.comp:68-74
// for the current handle-table depth (Renderer.cpp:598-720).  A HIP stream
// starts executing work when it is enqueued (a Vulkan command buffer only at
// submission), so the staging fill that the reference performs while
// recording the draws (StateSet.cpp:233-237) happens before the copy here;
// recordSceneRendering then only reports the per-StateSet draw ranges.
//
// Resident drawable records (on by default on a device): the drawable list
// stays in device memory between frames.  When the StateSet ranges are the
// frame before's, only the records that changed since (append, removal's
// move into the hole, in-place Drawable::create) are sent, as (slot, record)
// pairs applied by cadrScatterDrawables; when ranges moved, the list is
// rebuilt in a second device buffer by device-to-device copies of every
// StateSet it already holds plus those deltas, and StateSets new to the list
// are copied whole.  The list the gather reads is then the one the full copy
// would give.  setResidentDrawables(false) restores the reference's full
// staging memcpy and copy each frame (StateSet.cpp:233-237,
// Renderer.cpp:635-644).  drawableBufferAddress() can change from frame to
// frame in resident mode.
//
// Cull frame (an extension: the reference draws every drawable).
// recordDrawableCulling(stream, views, n), after recordDrawableProcessing,
// culls the frame's drawables for n views on the GPU, selects each one's LOD
// and compacts the survivors per StateSet and view:
//   - one cull item per matrix of each drawable's MatrixList (an instance),
//     in the draw-range order; its world sphere is the Geometry's bounding
//     sphere under that matrix (BoundingSphere.h:70-87), computed on the
//     device from the resident drawable list through the handle table;
//   - LOD k of a drawable is the PrimitiveSet record at its offset + 8k
//     (the Geometry's numLods() of them), chosen by projected radius;
//   - for view w, StateSet range k's visible items are compacted to the front
//     of [itemBase_k, itemBase_k + itemCount_k) of the view's
//     VkDrawIndexedIndirectCommand array (cullCommandBufferAddress() +
//     w * cullCommandStride() * 20), their number at
//     cullCountBufferAddress() + (w * numRanges + k) * 4: the operands of
//     vkCmdDrawIndexedIndirectCount;
//   - each command draws one instance: firstInstance is the matrix index in
//     the drawable's MatrixList (matrices[gl_InstanceIndex] in the reference's
//     shaders), and with setCullDrawablePointers(true) the drawable's
//     DrawablePointers record (from the gather) sits at the command's slot of
//     cullPointersBufferAddress(), for DrawablePointers[gl_DrawID]
//     (examples/RenderingPerformance/shader.vert:99-123).
// recordSceneRendering then reports itemBase / itemCount of every range.
//
// FrameInfo carries the reference's fields (FrameInfo.h:15-32); GPU
// timestamps are HIP events converted to nanoseconds from the frame's first
// GPU event.
#pragma once

#include "Export.h"

#include <cstdint>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "DataStorage.h"
#include "Drawable.h"
#include "../cadr_hip.h"

namespace CadR {

struct FrameInfo {
	size_t frameNumber = size_t(-1);
	uint64_t gpuBeginFrame = 0;
	uint64_t cpuBeginFrame = 0;
	uint64_t cpuPrepareRecordingBegin = 0;
	uint64_t cpuPrepareRecordingEnd = 0;
	uint64_t cpuRecordStateSetsBegin = 0;
	uint64_t cpuRecordStateSetsEnd = 0;
	std::vector<uint64_t> cpuUserTimestampList;
	uint64_t cpuEndFrame = 0;
	uint64_t gpuBeginExecution = 0;
	uint64_t gpuAfterTransfersAndBeforeDrawableProcessing = 0;
	uint64_t gpuAfterDrawableProcessingAndBeforeRendering = 0;
	uint64_t gpuEndExecution = 0;
	static constexpr uint32_t gpuTimestampPoolSize = 4;
};

class CADR_EXPORT Renderer {
public:
	static constexpr size_t smallMemorySize = DataStorage::smallMemorySize;
	static constexpr size_t mediumMemorySize = DataStorage::mediumMemorySize;
	static constexpr size_t largeMemorySize = DataStorage::largeMemorySize;
	static constexpr uint32_t drawablePointersRecordSize = 4 * sizeof(uint64_t);

	// device-less renderer: host memory, bookkeeping only (Renderer.cpp:64-74)
	explicit Renderer(bool makeDefault = true);
	// renderer on HIP device deviceIndex
	explicit Renderer(int deviceIndex, bool makeDefault = true);
	~Renderer();
	Renderer(const Renderer&) = delete;
	Renderer& operator=(const Renderer&) = delete;
	void finalize();

	static Renderer& get() { return *_defaultRenderer; }
	static void set(Renderer& r) noexcept { _defaultRenderer = &r; }

	// frame API
	size_t beginFrame();
	void executeCopyOperations();
	void beginRecording(void* stream);
	size_t prepareSceneRendering(StateSet& root);
	void recordDrawableProcessing(void* stream, size_t numDrawables);
	void recordSceneRendering(void* stream, StateSet& root, std::vector<DrawRange>& ranges);
	void endRecording(void* stream);
	void endFrame();

	// getters
	bool hasDevice() const { return _device >= 0; }
	int deviceIndex() const { return _device; }
	size_t frameNumber() const noexcept { return _frameNumber; }
	DataStorage& dataStorage() const { return *_dataStorage; }
	uint64_t drawableBufferAddress() const { return _drawableBuffer; }
	size_t drawableBufferSize() const { return _capacity * sizeof(DrawableGpuData); }
	DrawableGpuData* drawableStagingData() const { return _staging; }
	uint64_t drawIndirectBufferAddress() const { return _indirectBuffer; }
	uint64_t drawablePointersBufferAddress() const { return _pointersBuffer; }
	size_t bufferCapacity() const { return _capacity; }

	// resident drawable records (see above) and the host -> device bytes the
	// last recordDrawableProcessing sent for the drawable list
	void setResidentDrawables(bool on);
	bool residentDrawables() const { return _resident; }
	size_t lastDrawableUploadBytes() const { return _lastUploadBytes; }

	// cull frame (see above)
	void recordDrawableCulling(void* stream, const cadrView* views, size_t numViews);
	void setCullConfig(const cadrCullConfig& singleView, const cadrCullConfig& batched);
	void setCullDrawablePointers(bool on) { _cullPointersOn = on; }
	// the command, count and DrawablePointers buffers of the cull frame as
	// shared buffers a Vulkan renderer imports (cadrSharedBuffer*): each
	// exportCull*Fd returns a new dma-buf descriptor the caller owns;
	// cullBufferGeneration() changes whenever they are reallocated (growth),
	// and the renderer imports them again
	void setCullBuffersShared(bool on);
	bool cullBuffersShared() const { return _cullShared; }
	int exportCullCommandBufferFd() const;
	int exportCullCountBufferFd() const;
	int exportCullPointersBufferFd() const;
	uint64_t cullBufferGeneration() const { return _cullGeneration; }
	size_t cullCommandBufferSize() const { return _cull.commandsCap; }
	size_t cullCountBufferSize() const { return _cull.countsCap; }
	size_t cullPointersBufferSize() const { return _cullPointersOn ? _cull.pointersCap : 0; }
	uint64_t cullCommandBufferAddress() const { return _cull.commands; }
	uint64_t cullCountBufferAddress() const { return _cull.counts; }
	uint64_t cullPointersBufferAddress() const { return _cullPointersOn ? _cull.pointers : 0; }
	uint64_t cullSphereBufferAddress() const { return _cull.spheres; }
	uint64_t cullMeshTableAddress() const { return _cull.meshes; }
	size_t cullCommandStride() const { return _cullItems; }
	size_t numCullItems() const { return _cullItems; }
	size_t numCullViews() const { return _cullViews; }
	size_t numCullRanges() const { return _cullRanges.size(); }
	size_t numMeshes() const { return _meshKeys.size(); }
	// counts of the last cull frame, [view][range] (waits for the device)
	void downloadCullCounts(std::vector<uint32_t>& counts) const;
	// bookkeeping of the scene graph: the mesh index of a (Geometry,
	// PrimitiveSet offset) pair, a Geometry that goes away, and any change of
	// drawables or MatrixList sizes (the cull items are rebuilt)
	uint32_t meshIndexOf(const Geometry& g, uint32_t primitiveSetOffset);
	void forgetGeometry(const Geometry* g) noexcept;
	void sceneChanged() noexcept { ++_sceneEpoch; }

	// frame timing (Renderer.cpp:1025-1066)
	void setCollectFrameInfo(bool on) { _collectFrameInfo = on; }
	bool collectFrameInfo() const { return _collectFrameInfo; }
	const FrameInfo& getFrameInfo();
	static uint64_t getCpuTimestamp();
	static double cpuTimestampPeriod() { return 1e-9; }
	static double gpuTimestampPeriod() { return 1e-9; }

private:
	void releaseBuffers() noexcept;
	void recordEvent(int index, void* stream);
	void uploadFullList(void* stream, size_t numDrawables);
	void uploadResidentList(void* stream, const std::vector<DrawRange>& ranges);
	void reserveDelta(size_t count);
	void updateCullLayout(void* stream);
	void growCull(uint64_t& buffer, size_t& capacity, size_t bytes, const char* what);
	void growCullShared(uint64_t& buffer, size_t& capacity, cadrSharedBuffer*& shared, size_t bytes,
	                    const char* what);
	int exportCullFd(const cadrSharedBuffer* shared, const char* what) const;
	void releaseCull() noexcept;

	static Renderer* _defaultRenderer;
	int _device = -1;
	std::unique_ptr<DataStorage> _dataStorage;
	size_t _frameNumber = size_t(-1);
	size_t _capacity = 0;
	uint64_t _drawableBuffer = 0, _indirectBuffer = 0, _pointersBuffer = 0;
	DrawableGpuData* _staging = nullptr;
	std::vector<DrawRange> _pendingRanges;
	StateSet* _preparedRoot = nullptr;
	bool _collectFrameInfo = false;
	FrameInfo _inProgress, _completed;
	void* _events[FrameInfo::gpuTimestampPoolSize] = {};
	bool _eventRecorded[FrameInfo::gpuTimestampPoolSize] = {};
	size_t _infoFrame = size_t(-1);

	// resident drawable list
	bool _resident = true;
	bool _residentValid = false;             // _drawableBuffer holds _residentRanges' records
	uint64_t _drawableBufferSpare = 0;       // rebuild target when ranges move
	std::vector<DrawRange> _residentRanges;  // layout of _drawableBuffer
	uint8_t* _delta = nullptr;               // pinned: records (48 B each) then slots (uint32)
	uint64_t _deltaDevice = 0;               // device copy of _delta
	size_t _deltaCapacity = 0;               // records _delta holds
	void* _uploadDone = nullptr;             // host staging may be rewritten once this completes
	bool _uploadPending = false;
	void* _listFree = nullptr;               // the last gather has read the drawable list
	bool _listFreePending = false;
	size_t _lastUploadBytes = 0;

	// cull frame
	size_t _processedFrame = size_t(-1);     // frame of the last recordDrawableProcessing
	size_t _culledFrame = size_t(-1);        // frame of the last recordDrawableCulling
	// pinned host copies of the job, views and mesh sources, read by async
	// copies: rewritten only once _cullCopied (recorded after those copies)
	// completed
	uint8_t* _cullPinned = nullptr;
	size_t _cullPinnedCap = 0;
	void* _cullCopied = nullptr;
	bool _cullCopyPending = false;
	size_t _processedCount = 0;
	uint64_t _sceneEpoch = 1;
	uint64_t _cullLayoutEpoch = 0;          // _sceneEpoch the item layout was built for
	std::vector<DrawRange> _cullRanges;      // ranges the item layout was built for
	bool _cullIdentity = true;               // one item per drawable
	size_t _cullItems = 0, _cullViews = 0;
	uint32_t _cullTileSize = 0;
	uint32_t _cullNumTiles = 0;
	bool _cullPointersOn = false;
	cadrCullConfig _cullSingle{1024, 8, 0, 0}, _cullBatched{256, 8, 0, 0};
	struct CullBuffers {
		uint64_t spheres = 0, meshIndex = 0, itemBase = 0, itemSource = 0, rangeBase = 0, tiles = 0;
		uint64_t upload = 0;  // a frame's job, views and mesh sources, one copy
		uint64_t commands = 0, counts = 0, pointers = 0, meshes = 0;
		size_t spheresCap = 0, meshIndexCap = 0, itemBaseCap = 0, itemSourceCap = 0, rangeBaseCap = 0, tilesCap = 0;
		size_t uploadCap = 0, commandsCap = 0, countsCap = 0, pointersCap = 0, meshesCap = 0;
		cadrSharedBuffer *sharedCommands = nullptr, *sharedCounts = nullptr, *sharedPointers = nullptr;
	} _cull;
	bool _cullShared = false;
	uint64_t _cullGeneration = 0;
	cadrCullWorkspace* _cullWs = nullptr;
	uint64_t _cullWsWords = 0;
	std::map<std::pair<const Geometry*, uint32_t>, uint32_t> _meshIndex;
	std::vector<std::pair<const Geometry*, uint32_t>> _meshKeys;   // geometry null once it is gone
};

}  // namespace CadR
