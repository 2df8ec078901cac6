"""ctypes binding of libcadr_hip.so (the C ABI declared in include/cadr_hip.h).

The library is built in-tree by ``make -C cadr_amd/csrc`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CADR_HIP_LIBRARY selects another build of the library (same C ABI)
LIB_PATH = os.environ.get("CADR_HIP_LIBRARY") or os.path.join(_HERE, "libcadr_hip.so")

c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
c_float = ctypes.c_float

# ---------------------------------------------------------------- layouts
# Byte layouts of include/cadr_hip.h as numpy dtypes.
DRAWABLE_DTYPE = np.dtype([("vertexDataHandle", "<u8"), ("indexDataHandle", "<u8"),
                           ("matrixListHandle", "<u8"), ("drawableDataHandle", "<u8"),
                           ("primitiveSetHandle", "<u8"), ("primitiveSetOffset", "<u4"),
                           ("padding", "<u4")])
MESH_DTYPE = np.dtype([("center", "<f4", (3,)), ("radius", "<f4"), ("lodCount", "<u4"),
                       ("vertexOffset", "<i4"), ("reserved", "<u4", (2,)),
                       ("indexCount", "<u4", (4,)), ("firstIndex", "<u4", (4,))])
VIEW_DTYPE = np.dtype([("viewRow", "<f4", (3, 4)), ("nearZ", "<f4"), ("farZ", "<f4"),
                       ("sideX", "<f4", (2,)), ("sideY", "<f4", (2,)), ("lodScale", "<f4"),
                       ("lodThreshold", "<f4", (3,)), ("reserved", "<f4", (2,))])
TILE_DTYPE = np.dtype([("start", "<u4"), ("count", "<u4"), ("stateSetBase", "<u4"),
                       ("stateSetAndFlags", "<u4"), ("chainIndex", "<u4"), ("reserved", "<u4", (3,))])
JOB_DTYPE = np.dtype([("spheres", "<u8"), ("meshIndex", "<u8"), ("meshes", "<u8"), ("tiles", "<u8"),
                      ("views", "<u8"), ("commands", "<u8"), ("counts", "<u8"),
                      ("numDrawables", "<u4"), ("numTiles", "<u4"), ("numStateSets", "<u4"),
                      ("numViews", "<u4"), ("commandStride", "<u4"), ("numMeshes", "<u4")])
CMD_DTYPE = np.dtype([("indexCount", "<u4"), ("instanceCount", "<u4"), ("firstIndex", "<u4"),
                      ("vertexOffset", "<i4"), ("firstInstance", "<u4")])
INDIRECT_DTYPE = np.dtype([("vertexCount", "<u4"), ("instanceCount", "<u4"), ("firstVertex", "<u4"),
                           ("firstInstance", "<u4")])
POINTERS_DTYPE = np.dtype([("vertexDataPtr", "<u8"), ("indexDataPtr", "<u8"), ("matrixListPtr", "<u8"),
                           ("drawableDataPtr", "<u8")])
LEVEL_DTYPE = np.dtype([("localMatrices", "<u8"), ("parentIndex", "<u8"), ("parentWorld", "<u8"),
                        ("worldMatrices", "<u8"), ("count", "<u8"), ("meshIndex", "<u8"),
                        ("meshes", "<u8"), ("spheres", "<u8"), ("numMeshes", "<u4"),
                        ("parentCount", "<u4")])

assert DRAWABLE_DTYPE.itemsize == 48 and MESH_DTYPE.itemsize == 64 and VIEW_DTYPE.itemsize == 96
assert TILE_DTYPE.itemsize == 32 and JOB_DTYPE.itemsize == 80 and CMD_DTYPE.itemsize == 20
assert LEVEL_DTYPE.itemsize == 72
# cadrSynthLevel (include/cadr_synth.h)
SYNTH_LEVEL_DTYPE = np.dtype([("seed", "<u8"), ("count", "<u8"), ("permuteKey", "<u8"), ("localMatrices", "<u8"),
                              ("parentIndex", "<u8"), ("meshIndex", "<u8"), ("meshes", "<u8"), ("spheres", "<u8"),
                              ("stream", "<u4"), ("transform", "<u4"), ("translationRange", "<f4"),
                              ("fanout", "<u4"), ("meshStream", "<u4"), ("numMeshes", "<u4")])
assert SYNTH_LEVEL_DTYPE.itemsize == 88
# cadrCullTuneJob / cadrCullTuneResult (include/cadr_hip.h)
TUNE_JOB_DTYPE = np.dtype([("job", JOB_DTYPE), ("stateSetCounts", "<u8")])
TUNE_RESULT_DTYPE = np.dtype([("blockSize", "<u4"), ("itemsPerThread", "<u4"), ("workgroupsPerCU", "<u4"),
                              ("flags", "<u4"), ("milliseconds", "<f4"), ("agrees", "<u4")])
assert TUNE_JOB_DTYPE.itemsize == 88 and TUNE_RESULT_DTYPE.itemsize == 24
TUNE_SINGLE_VIEW = 1
SYNTH_ROTATION_TRANSLATION = 0
SYNTH_SCALE_TRANSLATION = 1

TILE_FIRST = 0x80000000
TILE_LAST = 0x40000000
TILE_STATESET_MASK = 0x3FFFFFFF
CULL_NO_TILE_SPHERES = 1
CULL_WHOLE_TILES = 2
CULL_PATH_VIEW, CULL_PATH_BATCH, CULL_PATH_FUSED_VIEW, CULL_PATH_FUSED_BATCH = 1, 2, 4, 8

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_OUT_OF_RESOURCES = 2
ERROR_TIMEOUT = 3
ERROR_DEVICE = 4
ERROR_NOT_READY = 5


class PushConstants(ctypes.Structure):
    """processDrawables.comp:68-74 push-constant block."""
    _fields_ = [("handleTable", c_u64), ("drawableList", c_u64), ("indirectData", c_u64),
                ("drawablePointers", c_u64)]


class CullConfig(ctypes.Structure):
    _fields_ = [("blockSize", c_u32), ("itemsPerThread", c_u32), ("workgroupsPerCU", c_u32),
                ("flags", c_u32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("cull", CullConfig), ("maxBatchViews", c_u32), ("maxInflightBatches", c_u32),
                ("outputBudgetBytes", c_u64), ("maxQueuedRequests", c_u32), ("reserved", c_u32)]


class ServerScene(ctypes.Structure):
    _fields_ = [("spheres", c_u64), ("meshIndex", c_u64), ("meshes", c_u64), ("stateSetCounts", c_void_p),
                ("numStateSets", c_u32), ("numMeshes", c_u32), ("numDrawables", c_u32), ("reserved", c_u32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("requestId", c_u64), ("sceneId", c_u32), ("numStateSets", c_u32), ("commands", c_u64),
                ("counts", c_u64), ("visible", c_u64), ("submitNs", c_u64), ("completeNs", c_u64),
                ("status", ctypes.c_int32), ("tag", c_u32)]


class LoadSample(ctypes.Structure):
    _fields_ = [("result", ServerResult)]


class LoadSpec(ctypes.Structure):
    _fields_ = [("requestsPerSecond", ctypes.c_double), ("durationSeconds", ctypes.c_double), ("seed", c_u64),
                ("numHotScenes", c_u32), ("hotFraction", c_float), ("views", c_void_p), ("numViews", c_u32),
                ("sampleEvery", c_u32), ("samples", c_void_p), ("sampleCapacity", c_u32), ("reserved", c_u32)]


class LoadStats(ctypes.Structure):
    _fields_ = [("submitted", c_u64), ("completed", c_u64), ("refused", c_u64), ("elapsedSeconds", ctypes.c_double),
                ("requestsPerSecond", ctypes.c_double), ("drawablesPerSecond", ctypes.c_double),
                ("p50Ms", ctypes.c_double), ("p99Ms", ctypes.c_double), ("p999Ms", ctypes.c_double),
                ("maxMs", ctypes.c_double), ("meanMs", ctypes.c_double), ("visible", c_u64), ("launches", c_u64),
                ("meanViewsPerLaunch", ctypes.c_double), ("sampled", c_u64), ("failed", c_u64),
                ("sceneBytesRead", c_u64)]


class CadrError(RuntimeError):
    """A cadrResult other than CADR_SUCCESS (src/CadR/Exceptions.h analogue)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


# Exported symbols and their C signatures (restype, argtypes).
SIGNATURES = {
    "cadrGetResultString": (ctypes.c_char_p, [c_int]),
    "cadrGetLastErrorMessage": (ctypes.c_char_p, []),
    "cadrSetDevice": (c_int, [c_int]),
    "cadrGetDeviceCount": (c_int, [ctypes.POINTER(c_int)]),
    "cadrGetComputeUnitCount": (c_int, [ctypes.POINTER(c_int)]),
    "cadrDeviceMalloc": (c_int, [c_u64, ctypes.POINTER(c_u64)]),
    "cadrDeviceFree": (c_int, [c_u64]),
    "cadrHostMallocPinned": (c_int, [c_u64, ctypes.POINTER(c_void_p)]),
    "cadrHostFreePinned": (c_int, [c_void_p]),
    "cadrMemcpyHtoD": (c_int, [c_u64, c_void_p, c_u64, c_void_p]),
    "cadrMemcpyDtoH": (c_int, [c_void_p, c_u64, c_u64, c_void_p]),
    "cadrMemcpyDtoD": (c_int, [c_u64, c_u64, c_u64, c_void_p]),
    "cadrFillU32": (c_int, [c_u64, c_u32, c_u64, c_void_p]),
    "cadrStreamCreate": (c_int, [ctypes.POINTER(c_void_p)]),
    "cadrStreamDestroy": (c_int, [c_void_p]),
    "cadrStreamSynchronize": (c_int, [c_void_p]),
    "cadrDeviceSynchronize": (c_int, []),
    "cadrEventCreate": (c_int, [ctypes.POINTER(c_void_p)]),
    "cadrEventDestroy": (c_int, [c_void_p]),
    "cadrEventRecord": (c_int, [c_void_p, c_void_p]),
    "cadrEventSynchronize": (c_int, [c_void_p]),
    "cadrEventQuery": (c_int, [c_void_p]),
    "cadrEventElapsedTime": (c_int, [ctypes.POINTER(c_float), c_void_p, c_void_p]),
    "cadrTimeNextLaunch": (c_int, [c_void_p, c_void_p]),
    "cadrProcessDrawables": (c_int, [ctypes.POINTER(PushConstants), c_u32, c_u64, c_void_p]),
    "cadrScatterDrawables": (c_int, [c_u64, c_u64, c_u64, c_u64, c_u32, c_void_p]),
    "cadrPropagateTransforms": (c_int, [c_void_p, c_void_p]),
    "cadrPropagateTransformsBatch": (c_int, [c_u64, c_void_p, c_u32, c_void_p]),
    "cadrTransformPlanCreate": (c_int, [c_void_p, c_u32, c_void_p, ctypes.POINTER(c_void_p)]),
    "cadrSharedBufferCreate": (c_int, [c_u64, ctypes.POINTER(c_void_p)]),
    "cadrSharedBufferImport": (c_int, [c_int, c_u64, ctypes.POINTER(c_void_p)]),
    "cadrSharedBufferAddress": (c_u64, [c_void_p]),
    "cadrSharedBufferSize": (c_u64, [c_void_p]),
    "cadrSharedBufferExportFd": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
    "cadrSharedBufferDestroy": (None, [c_void_p]),
    "cadrPropagateTransformsPlanned": (c_int, [c_void_p, c_void_p]),
    "cadrTransformPlanDestroy": (None, [c_void_p]),
    "cadrComputeWorldSpheres": (c_int, [c_u64, c_u64, c_u64, c_u32, c_u64, c_u64, c_void_p]),
    "cadrMakeView": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_float, c_float,
                             c_void_p, c_void_p]),
    "cadrBuildCullTiles": (c_int, [c_void_p, c_u32, c_u32, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    "cadrCullWorkspaceCreate": (c_int, [c_u64, c_u32, ctypes.POINTER(c_void_p)]),
    "cadrCullWorkspaceDestroy": (c_int, [c_void_p]),
    "cadrCullWorkspaceSetWaitLimit": (c_int, [c_void_p, c_u32]),
    "cadrCullWorkspaceCheck": (c_int, [c_void_p]),
    "cadrCullBatch": (c_int, [ctypes.POINTER(CullConfig), c_u64, c_void_p, c_u32, c_void_p, c_void_p]),
    "cadrCullBatchNoAlloc": (c_int, [ctypes.POINTER(CullConfig), c_u64, c_void_p, c_u32, c_void_p]),
    "cadrCullView": (c_int, [ctypes.POINTER(CullConfig), c_u64, c_void_p, c_void_p, c_void_p]),
    "cadrCullTileSize": (c_u32, [ctypes.POINTER(CullConfig)]),
    "cadrCullConfigPaths": (c_u32, [ctypes.POINTER(CullConfig)]),
    "cadrPropagateCullView": (c_int, [c_void_p, ctypes.POINTER(CullConfig), c_u64, c_void_p, c_void_p, c_void_p]),
    "cadrCompactDrawablePointers": (c_int, [c_u64, c_u64, c_u64, c_u32, c_u32, c_u32, c_u64, c_u64, c_u64,
                                            c_void_p]),
    "cadrPropagateCullBatch": (c_int, [c_u64, c_void_p, ctypes.POINTER(CullConfig), c_u64, c_void_p, c_u32,
                                       c_void_p, c_void_p]),
    "cadrServerCreate": (c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(c_void_p)]),
    "cadrServerDestroy": (c_int, [c_void_p]),
    "cadrServerAddScene": (c_int, [c_void_p, ctypes.POINTER(ServerScene), ctypes.POINTER(c_u32)]),
    "cadrServerStart": (c_int, [c_void_p]),
    "cadrServerSubmit": (c_int, [c_void_p, c_u32, c_void_p, ctypes.POINTER(c_u64)]),
    "cadrServerPoll": (c_int, [c_void_p, c_void_p, c_u32, ctypes.POINTER(c_u32)]),
    "cadrServerCopyCounts": (c_int, [c_void_p, c_u64, c_void_p, c_u32]),
    "cadrServerRelease": (c_int, [c_void_p, c_u64]),
    "cadrServerDrain": (c_int, [c_void_p, c_u32]),
    "cadrServerRunLoad": (c_int, [c_void_p, ctypes.POINTER(LoadSpec), ctypes.POINTER(LoadStats)]),
    "cadrCullAutotune": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_void_p, c_u32, c_void_p, c_u32,
                                 ctypes.POINTER(c_u32), ctypes.POINTER(CullConfig), c_void_p]),
    "cadrSynthesizeLevel": (c_int, [c_void_p, c_void_p]),
    "cadrSynthPermute": (c_u64, [c_u64, c_u64, c_u64]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads libcadr_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C cadr_amd/csrc); there is no fallback implementation")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != SUCCESS:
        lib = load()
        msg = lib.cadrGetLastErrorMessage().decode(errors="replace")
        raise CadrError(code, msg or lib.cadrGetResultString(code).decode())


def ptr(a: np.ndarray) -> int:
    """Host address of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
