/*
 * cadr_hip.h — C ABI of the MI355X (gfx950) drawable-processing pass.
 *
 * This is the drop-in boundary of libcadr_hip.so.  Every entry point takes
 * plain integers, device addresses (uint64_t, as Vulkan buffer device
 * addresses are in the reference) and sizes; nothing here uses C++ or torch
 * types.  Errors are returned as cadrResult codes, never thrown.  The detail
 * text of the last error of the calling thread is kept by
 * cadrGetLastErrorMessage().
 *
 * Reference interfaces replaced (paths relative to the Rendering-FIT/CADR
 * tree):
 *   - the compute pass   src/CadR/shaders/processDrawables.comp:92-113
 *     and its recording  src/CadR/Renderer.cpp:598-720
 *     (Renderer::recordDrawableProcessing, push constants :671-683,
 *     32768-split dispatch :684-692);
 *   - the CPU-side hierarchy propagation examples/gltfReader/main.cpp:1538-1576
 *     (m = parentMatrix * node.matrix, glm mat4*mat4
 *     3rdParty/GLM/glm/detail/type_mat4x4.inl:630-648);
 *   - the world bounding sphere src/CadR/BoundingSphere.h:70-87;
 *   - the per-StateSet draw ranges src/CadR/StateSet.cpp:198-268
 *     (StateSet k owns commands [base_k, base_k + n_k) of the indirect buffer).
 * Frustum culling, LOD selection and compaction have no reference
 * implementation; their definition is given with cadrCullView below.
 *
 * Thread safety: like the reference (src/CadR has no locks), calls that share
 * a stream or a cull workspace must be serialised by the caller.
 */
#ifndef CADR_HIP_H
#define CADR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Results.  The reference throws CadR::LogicError / OutOfResources / Timeout
 * (src/CadR/Exceptions.h:13-40) and vk::SystemError; across the C ABI these
 * become the codes below. */
typedef enum cadrResult {
	CADR_SUCCESS = 0,
	CADR_ERROR_INVALID_VALUE = 1,     /* LogicError: bad argument or layout */
	CADR_ERROR_OUT_OF_RESOURCES = 2,  /* OutOfResources: allocation failed  */
	CADR_ERROR_TIMEOUT = 3,           /* Timeout: a wait did not finish     */
	CADR_ERROR_DEVICE = 4,            /* HIP runtime error (vk::SystemError) */
	CADR_ERROR_NOT_READY = 5          /* query: work not finished yet       */
} cadrResult;

const char* cadrGetResultString(int result);
const char* cadrGetLastErrorMessage(void);

/* HIP stream handle (hipStream_t); 0 means the legacy default stream. */
typedef void* cadrStream;
/* HIP event handle (hipEvent_t). */
typedef void* cadrEvent;

/* ------------------------------------------------------------------------ */
/* Device, memory and streams.  Stand-ins for the Vulkan device, DataMemory
 * blocks and the staging path (src/CadR/Renderer.cpp:97-313, 946-999,
 * src/CadR/DataMemory.cpp:255-397). */
int cadrSetDevice(int deviceIndex);
int cadrGetDeviceCount(int* count);
/* Number of compute units of the current device (256 on MI355X). */
int cadrGetComputeUnitCount(int* count);
int cadrDeviceMalloc(uint64_t numBytes, uint64_t* deviceAddress);
int cadrDeviceFree(uint64_t deviceAddress);
int cadrHostMallocPinned(uint64_t numBytes, void** hostPointer);
int cadrHostFreePinned(void* hostPointer);
int cadrMemcpyHtoD(uint64_t dst, const void* src, uint64_t numBytes, cadrStream stream);
int cadrMemcpyDtoH(void* dst, uint64_t src, uint64_t numBytes, cadrStream stream);
int cadrMemcpyDtoD(uint64_t dst, uint64_t src, uint64_t numBytes, cadrStream stream);
/* Fills numBytes/4 uint32 words with value using a kernel (capturable in a
 * HIP graph, unlike hipMemsetAsync). numBytes must be a multiple of 4. */
int cadrFillU32(uint64_t dst, uint32_t value, uint64_t numBytes, cadrStream stream);
int cadrStreamCreate(cadrStream* stream);
int cadrStreamDestroy(cadrStream stream);
int cadrStreamSynchronize(cadrStream stream);
int cadrDeviceSynchronize(void);
int cadrEventCreate(cadrEvent* event);
int cadrEventDestroy(cadrEvent event);
int cadrEventRecord(cadrEvent event, cadrStream stream);
int cadrEventSynchronize(cadrEvent event);
/* CADR_SUCCESS when complete, CADR_ERROR_NOT_READY otherwise. */
int cadrEventQuery(cadrEvent event);
int cadrEventElapsedTime(float* milliseconds, cadrEvent start, cadrEvent stop);
/* Kernel timestamps without marker commands: the next entry point this thread
 * calls among cadrPropagateTransforms, cadrPropagateTransformsBatch,
 * cadrPropagateTransformsPlanned, cadrProcessDrawables, cadrCullView,
 * cadrPropagateCullView, cadrCullBatch, cadrPropagateCullBatch,
 * cadrCullBatchNoAlloc and cadrCompactDrawablePointers records start and stop
 * (either may be NULL) on the dispatch of its main kernel, so
 * cadrEventElapsedTime(start, stop) is that kernel's own duration, and the
 * stream carries no extra commands between kernels (an event recorded with
 * cadrEventRecord is a command of its own and adds a few microseconds to a
 * chain of kernels).  An entry point that launches nothing records both back
 * to back; one that fails before launching leaves them pending.  The events
 * stay pending until such a call, also one made inside another entry point
 * on this thread (cadrCullAutotune's first candidate, the facade's frames);
 * cadrTimeNextLaunch(NULL, NULL) clears them.  Not for use while the stream
 * is being captured into a HIP graph.
 * Plays the part of the reference's vkCmdWriteTimestamp pair around its
 * compute dispatch (src/CadR/Renderer.cpp:660-700; both written at once
 * when there is nothing to dispatch, :600-617). */
int cadrTimeNextLaunch(cadrEvent start, cadrEvent stop);

/* ------------------------------------------------------------------------ */
/* Byte layouts shared with the reference (all little endian).              */

/* DrawableGpuData, 48 B: src/CadR/Drawable.h:32-43, static_assert
 * src/CadR/Drawable.cpp:13, GLSL view processDrawables.comp:17-26. */
typedef struct cadrDrawableGpuData {
	uint64_t vertexDataHandle;
	uint64_t indexDataHandle;
	uint64_t matrixListHandle;
	uint64_t drawableDataHandle;
	uint64_t primitiveSetHandle;
	uint32_t primitiveSetOffset;   /* byte offset added to the PrimitiveSet address */
	uint32_t padding;
} cadrDrawableGpuData;

/* PrimitiveSet GPU record, 8 B: processDrawables.comp:29-33,
 * src/CadR/PrimitiveSet.h:12-15. */
typedef struct cadrPrimitiveSet {
	uint32_t count;
	uint32_t first;
} cadrPrimitiveSet;

/* MatrixList GPU header, 64 B, followed by column-major float mat4[]:
 * processDrawables.comp:35-41, src/CadR/MatrixList.h:54-59. */
typedef struct cadrMatrixListHeader {
	uint32_t numMatrices;
	uint32_t capacity;
	uint64_t padding[7];
} cadrMatrixListHeader;

/* VkDrawIndirectCommand, 16 B: processDrawables.comp:43-50. */
typedef struct cadrDrawIndirectCommand {
	uint32_t vertexCount;
	uint32_t instanceCount;
	uint32_t firstVertex;
	uint32_t firstInstance;
} cadrDrawIndirectCommand;

/* DrawablePointers, 32 B: processDrawables.comp:52-59,
 * Renderer::drawablePointersRecordSize src/CadR/Renderer.h:201. */
typedef struct cadrDrawablePointers {
	uint64_t vertexDataPtr;
	uint64_t indexDataPtr;
	uint64_t matrixListPtr;
	uint64_t drawableDataPtr;
} cadrDrawablePointers;

/* VkDrawIndexedIndirectCommand, 20 B: the compacted output of the cull pass. */
typedef struct cadrDrawIndexedIndirectCommand {
	uint32_t indexCount;
	uint32_t instanceCount;
	uint32_t firstIndex;
	int32_t  vertexOffset;
	uint32_t firstInstance;
} cadrDrawIndexedIndirectCommand;

/* Push-constant block of processDrawables.comp:68-74, filled by
 * Renderer::recordDrawableProcessing (src/CadR/Renderer.cpp:671-683). */
typedef struct cadrProcessDrawablesPushConstants {
	uint64_t handleTable;        /* root table device address (DataStorage::handleTableDeviceAddress) */
	uint64_t drawableList;       /* cadrDrawableGpuData[numDrawables], 16-B aligned */
	uint64_t indirectData;       /* cadrDrawIndirectCommand[numDrawables], 16-B aligned */
	uint64_t drawablePointers;   /* cadrDrawablePointers[numDrawables], 16-B aligned */
} cadrProcessDrawablesPushConstants;

/* Handle table geometry: src/CadR/HandleTable.h:33-35. */
#define CADR_HANDLES_PER_TABLE 2048u
#define CADR_HANDLE_LEVEL_SHIFT 11u
#define CADR_HANDLE_LEVEL_MASK 0x7ffu
/* Largest drawable count one call accepts: the reference asserts
 * numDrawables < 32768*32768 (src/CadR/Renderer.cpp:687). */
#define CADR_MAX_DRAWABLES (32768ull * 32768ull - 1ull)

/* ------------------------------------------------------------------------ */
/* Gather pass: processDrawables.comp:92-113, dispatched as
 * Renderer::recordDrawableProcessing does (src/CadR/Renderer.cpp:668-692).
 *
 * For every drawable i < numDrawables:
 *   d   = drawableList[i]
 *   ps  = lookup(d.primitiveSetHandle) + d.primitiveSetOffset
 *   ml  = lookup(d.matrixListHandle)
 *   indirectData[i]     = { ps.count, ml.numMatrices, ps.first, 0 }
 *   drawablePointers[i] = { lookup(vertex), lookup(index), ml, lookup(drawableData) }
 * with lookup() walking handleLevel (1, 2 or 3) tables of 2048 uint64
 * entries (processDrawables.comp:77-89).
 * Invalid handles whose table walk leaves the tables that exist resolve to
 * address 0 (the reference reads unowned memory there), and a PrimitiveSet
 * or MatrixList that resolves to 0 yields zero fields instead of a read
 * through a null address.  numDrawables == 0 is a no-op, as in the
 * reference (Renderer.cpp:600-620). */
int cadrProcessDrawables(const cadrProcessDrawablesPushConstants* pushConstants,
                         uint32_t handleLevel, uint64_t numDrawables, cadrStream stream);

/* Delta update of a device-resident drawable list (48-byte DrawableGpuData
 * records, listSlots of them): drawableList[slots[i]] = records[i] for
 * i < count; slots at or past listSlots are skipped.  records (16-byte
 * aligned) and slots (uint32) are device arrays.  It replaces the per-frame
 * copy of the whole staged list (StateSet.cpp:233-237 memcpy into staging,
 * Renderer.cpp:635-644 staging -> device copy) by the records that changed;
 * the facade's Renderer uses it when the drawable order is unchanged.  The
 * order of writes to one slot listed twice is unspecified. */
/* Device buffers shared with a Vulkan renderer through dma-buf file
 * descriptors (VK_KHR_external_memory_fd): the cull's command, count and
 * DrawablePointers outputs in memory the renderer's
 * vkCmdDrawIndexedIndirectCount and shaders read directly.  Replaces the
 * renderer-owned buffers of Renderer.cpp:446-595 read by gl_DrawID
 * (RenderingPerformance/shader.vert:99-123).
 *   Create: device memory allocated here (size rounded up to the allocation
 *     granularity, cadrSharedBufferSize gives it), mapped at
 *     cadrSharedBufferAddress; ExportFd returns a new descriptor for the
 *     renderer's vkAllocateMemory + VkImportMemoryFdInfoKHR (the caller owns
 *     it; a successful Vulkan import takes it over).
 *   Import: memory the renderer allocated and exported (vkGetMemoryFdKHR),
 *     mapped here; a successful import takes over fd.
 * Destroy unmaps and releases this side's reference. */
typedef struct cadrSharedBuffer cadrSharedBuffer;
int cadrSharedBufferCreate(uint64_t size, cadrSharedBuffer** buffer);
int cadrSharedBufferImport(int fd, uint64_t size, cadrSharedBuffer** buffer);
uint64_t cadrSharedBufferAddress(const cadrSharedBuffer* buffer);
uint64_t cadrSharedBufferSize(const cadrSharedBuffer* buffer);
int cadrSharedBufferExportFd(const cadrSharedBuffer* buffer, int* fd);
void cadrSharedBufferDestroy(cadrSharedBuffer* buffer);

/* DrawablePointers for the cull's compacted draws, the reference-layout
 * companion of its commands.  The reference's shaders read
 * DrawablePointersRef[gl_DrawID] (CadPL/UberShaderPoints.vert:38) from a
 * per-StateSet base pushed before each StateSet's draw
 * (StateSet.cpp:240-258), with records as processDrawables.comp:52-59 writes
 * them.  For view w and StateSet k (base_k = stateSetBase[k], n = counts[w*K
 * + k]): pointersOut[w*commandStride + base_k + j] =
 * pointersIn[commands[w*commandStride + base_k + j].firstInstance] for
 * j < n (zeros for an index at or past numDrawables), so drawing StateSet k
 * of view w with vkCmdDrawIndexedIndirectCount from slot base_k and
 * pushing pointersOut + (w*commandStride + base_k)*32 as the StateSet's
 * pointer base gives every draw its own drawable's record.  Slots past a
 * segment's count are left unwritten.  commands / counts as a cull job
 * writes them, stateSetBase uint32[K] (device), pointersIn / pointersOut
 * 32-byte records (16-byte aligned, device). */
int cadrCompactDrawablePointers(uint64_t commands, uint64_t counts, uint64_t stateSetBase, uint32_t numViews,
                                uint32_t numStateSets, uint32_t commandStride, uint64_t pointersIn,
                                uint64_t numDrawables, uint64_t pointersOut, cadrStream stream);

int cadrScatterDrawables(uint64_t drawableList, uint64_t listSlots, uint64_t records, uint64_t slots,
                         uint32_t count, cadrStream stream);

/* ------------------------------------------------------------------------ */
/* Transform propagation (one tree level per call) and world bounding
 * spheres.  Matrices are column-major float mat4 (64 B), as glm::mat4 and
 * the MatrixList payload.
 *
 *   world[i] = parentWorld[parentIndex[i]] * local[i]    (glm operator*)
 * evaluated column by column as
 *   world[i][j] = ((P[0]*L[j][0] + P[1]*L[j][1]) + P[2]*L[j][2]) + P[3]*L[j][3]
 * in float with no fused multiply-add (type_mat4x4.inl:643-646).  A level
 * with parentWorld == 0 is a root level: world = local, copied bit for bit.
 *
 * If spheres != 0, the level also writes the world bounding sphere of each
 * node (BoundingSphere.h:70-87) from the mesh of meshIndex[i]:
 *   c' = mat3(W)*c + W[3].xyz     (mat3*vec3 left to right, type_mat3x3.inl:470-473)
 *   r' = sqrt(max(max(|W0|^2, |W1|^2), |W2|^2)) * r,  |v|^2 = (x*x + y*y) + z*z
 * as float4 {x, y, z, r}. */
typedef struct cadrMesh {
	float    center[3];         /* local bounding sphere                     */
	float    radius;
	uint32_t lodCount;          /* 1..4; LOD k is PrimitiveSet k of the mesh */
	int32_t  vertexOffset;      /* copied into the indexed draw command      */
	uint32_t reserved[2];
	uint32_t indexCount[4];     /* per LOD {count, first} (PrimitiveSet.h)   */
	uint32_t firstIndex[4];
} cadrMesh;                     /* 64 B */

typedef struct cadrTransformLevel {
	uint64_t localMatrices;     /* float mat4[count]                          */
	uint64_t parentIndex;       /* uint32[count]; ignored for a root level     */
	uint64_t parentWorld;       /* float mat4[] of the previous level, or 0    */
	uint64_t worldMatrices;     /* float mat4[count] output                    */
	uint64_t count;
	uint64_t meshIndex;         /* uint32[count]; used when spheres != 0       */
	uint64_t meshes;            /* cadrMesh[]                                  */
	uint64_t spheres;           /* float4[count] output, or 0                  */
	uint32_t numMeshes;         /* meshIndex >= numMeshes gives BoundingSphere::empty(), radius -inf */
	uint32_t parentCount;       /* nodes in the parent level; parentIndex >= parentCount acts as a root */
} cadrTransformLevel;

int cadrPropagateTransforms(const cadrTransformLevel* level, cadrStream stream);

/* Many independent levels (for example depth d of many trees) in one launch:
 * each computed exactly as cadrPropagateTransforms computes it.  deviceLevels
 * holds cadrTransformLevel[numLevels] in device memory, hostLevels the same
 * descriptors on the host (for checks).  No level may read another level's
 * world matrices in the same launch; dependent levels go to successive
 * launches on a stream.  A root level's world matrices are its local ones,
 * so a level may name a root level's localMatrices as its parentWorld and
 * then share that root level's launch.  numLevels <= 1024.  Capturable in a
 * HIP graph. */
int cadrPropagateTransformsBatch(uint64_t deviceLevels, const cadrTransformLevel* hostLevels, uint32_t numLevels,
                                 cadrStream stream);

/* A fixed batch of independent levels (the rules of
 * cadrPropagateTransformsBatch, any number of levels) whose 64-node work
 * units are ordered by the parent of their first node: the units in flight
 * at any time read nearby parent matrices, which then stay in the L2 where
 * many levels share one parent array and each level's nodes are ordered by
 * parent (the instances of a mesh gathered from a scene graph, each set of
 * instances a level).  Results are those of cadrPropagateTransformsBatch;
 * only the order of the work differs.  Create reads the levels' parent
 * indices (synchronising the stream); the plan keeps copies of the
 * descriptors, so the arrays they name must stay in place while it is used.
 * Replaces the per-parent recursion of gltfReader/main.cpp:1538-1576 for
 * instanced MatrixLists; no reference counterpart of its own. */
typedef struct cadrTransformPlan cadrTransformPlan;
int cadrTransformPlanCreate(const cadrTransformLevel* hostLevels, uint32_t numLevels, cadrStream stream,
                            cadrTransformPlan** plan);
int cadrPropagateTransformsPlanned(const cadrTransformPlan* plan, cadrStream stream);
void cadrTransformPlanDestroy(cadrTransformPlan* plan);

/* World spheres of drawables whose world matrices are already known (flat
 * scenes): spheres[i] = worldMatrices[i] * meshes[meshIndex[i]].sphere. */
int cadrComputeWorldSpheres(uint64_t worldMatrices, uint64_t meshIndex, uint64_t meshes,
                            uint32_t numMeshes, uint64_t count, uint64_t spheres, cadrStream stream);

/* ------------------------------------------------------------------------ */
/* Frustum cull, LOD selection and per-StateSet compaction.
 *
 * Scene: numDrawables drawables grouped by StateSet in the order the
 * reference assigns draw ranges (StateSet DFS, StateSet.cpp:198-268):
 * StateSet k owns drawables [base_k, base_k + n_k), base_k = n_0 + ... + n_{k-1}.
 * Per drawable: a world sphere float4 {x,y,z,r} and a uint32 mesh index.
 *
 * View: a world-to-view matrix in glm lookAtLH convention (+z forward) and
 * frustum constants, all precomputed on the host (cadrMakeView).  For each
 * drawable with sphere (c, r) and view V, in float without FMA:
 *   v.x = (V[0][0]*c.x + V[1][0]*c.y) + (V[2][0]*c.z + V[3][0])   (glm mat4*vec4
 *   v.y, v.z likewise                                                order, :561-572)
 *   visible = (v.z + r >= nearZ) && (v.z - r <= farZ)
 *          && (|v.x|*sideX[0] - v.z*sideX[1] <= r)
 *          && (|v.y|*sideY[0] - v.z*sideY[1] <= r)
 *   p   = r * lodScale                         (projected radius = p / v.z px)
 *   lod = (p < lodThreshold[0]*v.z) + (p < lodThreshold[1]*v.z) + (p < lodThreshold[2]*v.z)
 *   lod = min(lod, mesh.lodCount - 1)
 * Output for view w: commands_w[base_k + j], j = 0 .. count_k - 1, are the
 * visible drawables of StateSet k in increasing drawable order, each
 *   { mesh.indexCount[lod], 1, mesh.firstIndex[lod], mesh.vertexOffset, i }
 * (firstInstance = drawable index i), and counts_w[k] = count_k.
 * Command slots past count_k are left unwritten. */
typedef struct cadrView {
	float viewRow[3][4];        /* rows 0..2 of the world->view mat4: {V[0][r], V[1][r], V[2][r], V[3][r]} */
	float nearZ;
	float farZ;
	float sideX[2];             /* {cos, sin} of the horizontal half angle */
	float sideY[2];             /* {cos, sin} of the vertical half angle   */
	float lodScale;             /* viewportHeight/2 * P[1][1] of perspectiveLH_ZO */
	float lodThreshold[3];      /* projected-radius px thresholds, descending; +inf = unused */
	float reserved[2];
} cadrView;                     /* 96 B, 16-byte aligned in device arrays */

/* Builds a cadrView for a camera at eye looking along dir with the given up
 * vector, vertical field of view fovy (radians), aspect = width/height,
 * depth range [zNear, zFar] and viewport height in pixels.  The view
 * matrix follows glm::lookAtLH (ext/matrix_transform.inl), P[1][1] =
 * 1/tan(fovy/2) follows glm::perspectiveLH_ZO (ext/matrix_clip_space.inl:255-268).
 * lodThresholds (3 floats, descending) may be NULL for {64, 8, 1} px. */
int cadrMakeView(const float eye[3], const float dir[3], const float up[3], float fovy,
                 float aspect, float zNear, float zFar, float viewportHeight,
                 const float* lodThresholds, cadrView* out);

/* A cull tile: a run of at most tileSize drawables inside one StateSet.
 * Tiles are processed in array order (the ticket order); chainIndex numbers
 * the tiles StateSet by StateSet (all tiles of StateSet 0, then of StateSet 1,
 * ...) and indexes the look-back status words, so the predecessor of a tile in
 * its StateSet is the tile with chainIndex - 1, which must come earlier in the
 * array. */
typedef struct cadrCullTile {
	uint32_t start;             /* first drawable                                  */
	uint32_t count;             /* 0 only for the single tile of an empty StateSet */
	uint32_t stateSetBase;      /* base_k of the tile's StateSet                   */
	uint32_t stateSetAndFlags;  /* k | CADR_TILE_FIRST | CADR_TILE_LAST            */
	uint32_t chainIndex;        /* StateSet-major position of the tile             */
	uint32_t reserved[3];
} cadrCullTile;                 /* 32 B */
#define CADR_TILE_FIRST 0x80000000u
#define CADR_TILE_LAST  0x40000000u
#define CADR_TILE_STATESET_MASK 0x3fffffffu

/* Splits StateSets of sizes stateSetCounts[0..numStateSets) into tiles of at
 * most tileSize drawables, emitted round-robin over the StateSets (the q-th
 * tile of every StateSet, then the (q+1)-th, ...): consecutive tiles of one
 * StateSet are numStateSets tickets apart, which keeps look-back walks short.
 * With tiles == NULL it only counts them.  A StateSet may hold at most
 * CADR_TILE_STATESET_MASK (2^30 - 1) drawables (compaction offsets travel as
 * 30-bit values) and a scene at most 2^32 - 1; larger counts are refused
 * with CADR_ERROR_INVALID_VALUE. */
int cadrBuildCullTiles(const uint32_t* stateSetCounts, uint32_t numStateSets, uint32_t tileSize,
                       cadrCullTile* tiles, uint32_t capacity, uint32_t* numTiles);

/* One scene culled for a run of views: the unit of a batched launch. */
typedef struct cadrCullJob {
	uint64_t spheres;           /* float4[numDrawables]                        */
	uint64_t meshIndex;         /* uint32[numDrawables]                        */
	uint64_t meshes;            /* cadrMesh[]                                  */
	uint64_t tiles;             /* cadrCullTile[numTiles] from cadrBuildCullTiles for the
	                             * launch's tile size (cadrCullTileSize); a tile longer
	                             * than that is processed only up to it */
	uint64_t views;             /* cadrView[numViews]                          */
	uint64_t commands;          /* cadrDrawIndexedIndirectCommand[numViews][commandStride] */
	uint64_t counts;            /* uint32[numViews][numStateSets]              */
	uint32_t numDrawables;
	uint32_t numTiles;
	uint32_t numStateSets;
	uint32_t numViews;
	uint32_t commandStride;     /* commands per view, >= numDrawables          */
	uint32_t numMeshes;         /* drawables with meshIndex >= numMeshes are never visible */
} cadrCullJob;                  /* 80 B */

/* Kernel shape.  tileSize = blockSize * itemsPerThread. */
typedef struct cadrCullConfig {
	uint32_t blockSize;         /* 256 or 512 threads                          */
	uint32_t itemsPerThread;    /* 2, 4, 8 or 16                               */
	uint32_t workgroupsPerCU;   /* persistent grid = CUs * this (0 = default)  */
	uint32_t flags;             /* CADR_CULL_* below                           */
} cadrCullConfig;
/* Test every drawable (no per-tile bounding sphere classification). */
#define CADR_CULL_NO_TILE_SPHERES 1u
/* Batched launches: every work item is one tile for all of its job's views.
 * By default the last (persistent grid size) tiles of a job with more than
 * 64 views are handed out as one work item per (tile, 64-view group), so the
 * launch's last work items are short and its workgroups finish together. */
#define CADR_CULL_WHOLE_TILES 2u

/* Device workspace of the cull launches: look-back status words, tile
 * tickets, the launch epoch and per-job offsets.  One workspace serves one
 * stream at a time; a launch needs statusWords >= sum over its jobs of
 * numViews * numTiles and numJobs <= maxJobs.  Status words carry the launch
 * epoch, so nothing is cleared between launches. */
typedef struct cadrCullWorkspace cadrCullWorkspace;
int cadrCullWorkspaceCreate(uint64_t statusWords, uint32_t maxJobs, cadrCullWorkspace** workspace);
int cadrCullWorkspaceDestroy(cadrCullWorkspace* workspace);
/* Look-back waits are bounded, so a launch ends whatever tile list it is
 * given.  A well-formed list (cadrBuildCullTiles) only waits on tiles with
 * earlier tickets, which are running or done; a caller-built list that names
 * a predecessor no tile publishes would otherwise wait forever.  A wait that
 * lasts longer than the limit (default 1 s, at least 1 us) gives up, every
 * other wait of the same launch gives up at its next unready poll, and the
 * launch's outputs are unspecified.  cadrCullWorkspaceCheck waits for the
 * device and returns CADR_ERROR_TIMEOUT (CadR::Timeout) if a launch on the
 * workspace gave up since the previous check, else CADR_SUCCESS; it clears
 * the record. */
int cadrCullWorkspaceSetWaitLimit(cadrCullWorkspace* workspace, uint32_t microseconds);
int cadrCullWorkspaceCheck(cadrCullWorkspace* workspace);

/* Culls numJobs jobs (job descriptors in device memory, cadrCullJob[numJobs])
 * in one persistent launch.  hostJobs is a host copy of the same
 * descriptors, used for argument checks and sizing; it is not read by the
 * device.  Capturable in a HIP graph. */
int cadrCullBatch(const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                  uint32_t numJobs, cadrCullWorkspace* workspace, cadrStream stream);

/* cadrCullBatch without a workspace and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its workspace is a fixed array in the library's device code,
 * one per GPU, held until the process ends: at most 2^21 look-back status
 * words (the sum over jobs of numViews x numTiles) and 4096 jobs per call;
 * larger calls are refused with CADR_ERROR_INVALID_VALUE.  The first call on
 * a device clears that device's array and waits for the clear.  Results are
 * bit-identical to cadrCullBatch with the same config and jobs.  All calls on
 * one GPU share its array: calls that overlap on one GPU (two streams, two
 * threads) race and must be serialised by the caller.  Look-back waits are
 * bounded as for a workspace; a give-up is not reported. */
int cadrCullBatchNoAlloc(const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                         uint32_t numJobs, cadrStream stream);

/* One scene, one view: cadrCullBatch of a single job with numViews == 1. */
int cadrCullView(const cadrCullConfig* config, uint64_t deviceJob, const cadrCullJob* hostJob,
                 cadrCullWorkspace* workspace, cadrStream stream);

/* The last transform level fused with the one-view cull, in one launch:
 * cadrPropagateTransforms(leaf) followed by cadrCullView(job), computed
 * tile by tile so that the world spheres never travel through HBM (they are
 * still written when leaf->spheres != 0).  Node i of the leaf level is
 * drawable i of the job (leaf->count == numDrawables); the spheres are made
 * from the job's meshIndex and meshes (the leaf's own meshIndex/meshes are
 * not read).  World matrices, spheres, commands and counts are bit-identical
 * to the two separate calls.  config (blockSize x itemsPerThread): 256 x 2/4/8/16,
 * 512 x 2/4 or 1024 x 1/2 (the last two shapes exist for this entry point only). */
int cadrPropagateCullView(const cadrTransformLevel* leaf, const cadrCullConfig* config, uint64_t deviceJob,
                          const cadrCullJob* hostJob, cadrCullWorkspace* workspace, cadrStream stream);

/* The leaf level of every job fused with the batched cull, in one persistent
 * launch: cadrPropagateTransformsBatch(leaves) followed by cadrCullBatch(jobs),
 * tile by tile, each tile's world matrices computed once for all of its job's
 * views (the propagation of examples/gltfReader/main.cpp:1538-1576 feeding the
 * per-frame pass that Renderer::recordDrawableProcessing, Renderer.cpp:598-720,
 * records).  deviceLeaves holds cadrTransformLevel[numJobs] in device memory,
 * hostLeaves the same descriptors on the host (for checks); leaf j describes
 * job j exactly as the leaf of cadrPropagateCullView does (leaf count ==
 * numDrawables, spheres from the job's meshIndex and meshes, written only when
 * the leaf's spheres != 0).  World matrices are written for the drawables the
 * job's tiles cover (all of them for tiles from cadrBuildCullTiles at this
 * tile size).  Results are bit-identical to the two separate calls.  config:
 * 256 x 4/8 or 512 x 4.  Capturable in a HIP graph. */
int cadrPropagateCullBatch(uint64_t deviceLeaves, const cadrTransformLevel* hostLeaves,
                           const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                           uint32_t numJobs, cadrCullWorkspace* workspace, cadrStream stream);

/* Kernel tile size for a config (blockSize * itemsPerThread), 0 if unsupported. */
uint32_t cadrCullTileSize(const cadrCullConfig* config);

/* The entry points that have a kernel of this shape (0: none):
 * cadrCullView (which runs the batched kernel when the shape has no one-view
 * kernel), cadrCullBatch, cadrPropagateCullView, cadrPropagateCullBatch. */
#define CADR_CULL_PATH_VIEW 1u
#define CADR_CULL_PATH_BATCH 2u
#define CADR_CULL_PATH_FUSED_VIEW 4u
#define CADR_CULL_PATH_FUSED_BATCH 8u
uint32_t cadrCullConfigPaths(const cadrCullConfig* config);

/* Autotuner of the cull launch shape (workgroup size, drawables per thread,
 * hence the scan tile, tile-sphere classification, workgroups per CU).
 * Times every candidate on the caller's own jobs — real launches that write
 * the jobs' commands and counts — with HIP events over `reps` launches after
 * one warm-up launch, and returns the fastest in *best.  Tiles are built
 * per candidate tile size from stateSetCounts (the jobs' `tiles` fields are
 * ignored).  With candidates == NULL every supported shape is tried (with
 * and without tile spheres for batched launches).  results (optional)
 * receives one entry per candidate; `agrees` is 1 when the candidate's
 * counts equal those of the first candidate that ran.
 * CADR_TUNE_SINGLE_VIEW times cadrCullView (one job with one view).
 * No counterpart in the reference. */
typedef struct cadrCullTuneJob {
	cadrCullJob job;                /* tiles ignored                          */
	const uint32_t* stateSetCounts; /* host, job.numStateSets entries         */
} cadrCullTuneJob;

typedef struct cadrCullTuneResult {
	cadrCullConfig config;
	float milliseconds;             /* mean per launch; +inf if it failed     */
	uint32_t agrees;
} cadrCullTuneResult;
#define CADR_TUNE_SINGLE_VIEW 1u

int cadrCullAutotune(const cadrCullTuneJob* jobs, uint32_t numJobs, uint32_t flags, uint32_t reps,
                     const cadrCullConfig* candidates, uint32_t numCandidates,
                     cadrCullTuneResult* results, uint32_t resultCapacity, uint32_t* numResults,
                     cadrCullConfig* best, cadrStream stream);

/* ------------------------------------------------------------------------ */
/* Cull inputs from the reference's own records, and the reference's draw
 * contract for the compacted commands (the CadR facade's cull frame,
 * Renderer::recordDrawableCulling).  The reference binds per drawable a
 * Geometry's PrimitiveSet allocation at a byte offset and a MatrixList
 * (DrawableGpuData, src/CadR/Drawable.h:32-43), and its consumer reads
 * DrawablePointers[gl_DrawID] and matrices[gl_InstanceIndex]
 * (examples/RenderingPerformance/shader.vert:99-123).  No reference
 * counterpart: the reference draws every drawable.
 *
 * A mesh of the cull's table is one (Geometry, PrimitiveSet offset) pair with
 * its local bounding sphere and LOD count; LOD k is the PrimitiveSet record
 * {count, first} at lookup(primitiveSetHandle) + primitiveSetOffset + 8k,
 * read as processDrawables.comp:97 reads LOD 0. */
typedef struct cadrMeshSource {
	uint64_t primitiveSetHandle;
	uint32_t primitiveSetOffset;
	uint32_t lodCount;             /* 1..4 (larger is 4); 0: the mesh no longer exists */
	float    sphere[4];            /* local bounding sphere {centre xyz, radius}       */
} cadrMeshSource;                  /* 32 B */

/* meshes[m] = the cadrMesh of sources[m]: its sphere and lodCount,
 * vertexOffset 0, indexCount / firstIndex of each LOD < lodCount (zero
 * beyond).  A source with lodCount 0, or whose PrimitiveSet handle resolves
 * to 0, gets the empty sphere (radius -inf) and one zero LOD: its drawables
 * are never visible.  sources and meshes are device arrays (16-byte aligned). */
int cadrBuildMeshTable(uint64_t handleTable, uint32_t handleLevel, uint64_t sources, uint32_t count,
                       uint64_t meshes, cadrStream stream);

/* Cull items of a device drawable list.  Item `it` stands for matrix j of
 * drawable i's MatrixList (one instance): with itemBase, i = itemDrawable[it]
 * and j = it - itemBase[i] (the items of a drawable are consecutive); without
 * it, item i is matrix 0 of drawable i.
 *   spheres[it]   = world sphere of meshes[d.padding] under matrix j
 *                   (BoundingSphere.h:70-87, as cadrComputeWorldSpheres), or
 *                   the empty sphere {0, 0, 0, -inf} when j >= numMatrices of
 *                   the list, the MatrixList handle resolves to 0,
 *                   d.padding >= numMeshes or i >= numDrawables;
 *   meshIndex[it] = d.padding (the mesh index, kept by the facade in the
 *                   record's padding word), ~0 for i >= numDrawables. */
typedef struct cadrCullItemsDesc {
	uint64_t handleTable;          /* root table (DataStorage::handleTableDeviceAddress) */
	uint64_t drawableList;         /* cadrDrawableGpuData[numDrawables]                  */
	uint64_t itemBase;             /* uint32[numDrawables], or 0: one item per drawable  */
	uint64_t itemDrawable;         /* uint32[numItems] (with itemBase), or 0             */
	uint64_t meshes;               /* cadrMesh[numMeshes]                                */
	uint64_t spheres;              /* float4[numItems], written                          */
	uint64_t meshIndex;            /* uint32[numItems], written                          */
	uint64_t numDrawables;
	uint64_t numItems;
	uint32_t handleLevel;
	uint32_t numMeshes;
} cadrCullItemsDesc;

int cadrBuildCullItems(const cadrCullItemsDesc* desc, cadrStream stream);

/* The reference's draw contract for compacted commands of cull items.  For
 * view w and StateSet k (base_k = stateSetBase[k], n = counts[w*K + k]) and
 * j < n, with c = commands[w*commandStride + base_k + j], item = c.firstInstance,
 * i = itemDrawable[item] and m = item - itemBase[i] (i = item, m = 0 without
 * itemBase):
 *   c.firstInstance = m      (gl_InstanceIndex = m reads the item's matrix)
 *   pointersOut[w*commandStride + base_k + j] = pointersIn[i]
 *                            (the drawable's DrawablePointers, from
 *                            cadrProcessDrawables; only when pointersOut != 0;
 *                            zeros for i >= numDrawables)
 * A command whose item index is at or past numItems is left as it is. */
int cadrFinishCullCommands(uint64_t commands, uint64_t counts, uint64_t stateSetBase, uint32_t numViews,
                           uint32_t numStateSets, uint32_t commandStride, uint64_t itemBase, uint64_t itemDrawable,
                           uint64_t numItems, uint64_t pointersIn, uint64_t numDrawables, uint64_t pointersOut,
                           cadrStream stream);

/* ------------------------------------------------------------------------ */
/* Serving: scenes resident in HBM, cull requests (scene, view) admitted by
 * output memory and batch slots, and batched continuously — every launch
 * takes the requests that are waiting when a slot frees, groups them by
 * scene into jobs of one persistent cull launch, and several launches run
 * at once on separate streams.  A request's result is exactly what
 * cadrCullView computes for its view alone (same kernel arithmetic, same
 * compaction order), whatever it was batched with.
 * The reference has no serving mode; this is an extension of the path. */
typedef struct cadrServer cadrServer;

typedef struct cadrServerConfig {
	cadrCullConfig cull;            /* kernel shape of every launch               */
	uint32_t maxBatchViews;         /* views per launch (e.g. 256)                */
	uint32_t maxInflightBatches;    /* launches in flight, one stream each        */
	uint64_t outputBudgetBytes;     /* device memory for request outputs          */
	uint32_t maxQueuedRequests;     /* submissions beyond this are refused        */
	uint32_t reserved;
} cadrServerConfig;

typedef struct cadrServerScene {
	uint64_t spheres;               /* float4[numDrawables], device, owned by the caller */
	uint64_t meshIndex;             /* uint32[numDrawables], device              */
	uint64_t meshes;                /* cadrMesh[numMeshes], device               */
	const uint32_t* stateSetCounts; /* host, numStateSets entries                */
	uint32_t numStateSets;
	uint32_t numMeshes;
	uint32_t numDrawables;
	uint32_t reserved;
} cadrServerScene;

typedef struct cadrServerResult {
	uint64_t requestId;
	uint32_t sceneId;
	uint32_t numStateSets;
	uint64_t commands;              /* device cadrDrawIndexedIndirectCommand[numDrawables], StateSet ranges */
	uint64_t counts;                /* device uint32[numStateSets]               */
	uint64_t visible;               /* sum of the counts                         */
	uint64_t submitNs;              /* steady-clock times of submission and of   */
	uint64_t completeNs;            /* the counts being visible on the host      */
	int32_t status;                 /* CADR_SUCCESS, or the error of the request's launch
	                                 * (its outputs are then unspecified, counts 0) */
	uint32_t tag;                   /* cadrServerRunLoad: index of the request's view; 0 otherwise */
} cadrServerResult;

int cadrServerCreate(const cadrServerConfig* config, cadrServer** server);
int cadrServerDestroy(cadrServer* server);
/* Scenes are added before cadrServerStart; their arrays stay owned by the caller. */
int cadrServerAddScene(cadrServer* server, const cadrServerScene* scene, uint32_t* sceneId);
int cadrServerStart(cadrServer* server);
/* Thread-safe.  CADR_ERROR_OUT_OF_RESOURCES when the queue is full or the
 * request's output can never fit the budget. */
int cadrServerSubmit(cadrServer* server, uint32_t sceneId, const cadrView* view, uint64_t* requestId);
/* Completed requests, oldest first; count may be 0.  A request whose launch
 * failed on the device completes with a non-zero status; the message of the
 * first such failure is kept by cadrGetLastErrorMessage of the polling
 * thread when that request is returned. */
int cadrServerPoll(cadrServer* server, cadrServerResult* results, uint32_t capacity, uint32_t* count);
/* Host copy of a completed request's per-StateSet counts. */
int cadrServerCopyCounts(cadrServer* server, uint64_t requestId, uint32_t* counts, uint32_t capacity);
/* Frees a completed request's output memory (its device arrays become invalid). */
int cadrServerRelease(cadrServer* server, uint64_t requestId);
/* Waits until no request is queued or running, or timeoutMs passes (CADR_ERROR_TIMEOUT). */
int cadrServerDrain(cadrServer* server, uint32_t timeoutMs);

/* Open-loop load: Poisson arrivals at requestsPerSecond for durationSeconds;
 * hotFraction of the requests go uniformly to scenes [0, numHotScenes), the
 * rest uniformly to the others; each request takes a view uniformly from
 * views[0..numViews).  Completed requests are released at once, except a
 * sample: with sampleEvery > 0, every sampleEvery-th completion (in
 * completion order, up to sampleCapacity) is kept in samples[] with its
 * outputs still allocated, for the caller to check and then free with
 * cadrServerRelease.  Kept outputs count against the output budget. */
typedef struct cadrLoadSample {
	cadrServerResult result;        /* result.tag = index of the request's view */
} cadrLoadSample;

typedef struct cadrLoadSpec {
	double requestsPerSecond;
	double durationSeconds;
	uint64_t seed;
	uint32_t numHotScenes;
	float hotFraction;
	const cadrView* views;          /* host */
	uint32_t numViews;
	uint32_t sampleEvery;           /* 0: keep no samples                        */
	cadrLoadSample* samples;        /* host, sampleCapacity entries, or NULL     */
	uint32_t sampleCapacity;
	uint32_t reserved;
} cadrLoadSpec;

typedef struct cadrLoadStats {
	uint64_t submitted, completed, refused;
	double elapsedSeconds;          /* first arrival to last completion          */
	double requestsPerSecond;       /* completed / elapsed                       */
	double drawablesPerSecond;      /* culled (drawable, view) tests / elapsed   */
	double p50Ms, p99Ms, p999Ms, maxMs, meanMs;
	uint64_t visible;
	uint64_t launches;
	double meanViewsPerLaunch;
	uint64_t sampled;               /* entries written to spec->samples          */
	uint64_t failed;                /* completed with a non-zero status          */
	uint64_t sceneBytesRead;        /* scene arrays the run's launches read: 20 B
	                                 * per drawable per job (a scene's drawables
	                                 * are read once for all its views in a launch) */
} cadrLoadStats;

int cadrServerRunLoad(cadrServer* server, const cadrLoadSpec* spec, cadrLoadStats* stats);

#ifdef __cplusplus
}
#endif

#endif /* CADR_HIP_H */
