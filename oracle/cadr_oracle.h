/*
 * CPU oracle of the drawable-processing pass — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the parity checker of libcadr_hip.so.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product never links, imports or calls it.
 *
 * It restates, in plain C compiled with -ffp-contract=off and no fast-math:
 *   - the gather of src/CadR/shaders/processDrawables.comp:77-113 over a
 *     host image of device memory (regions of {device base, size, bytes});
 *   - glm 0.9.9 mat4*mat4 (3rdParty/GLM/glm/detail/type_mat4x4.inl:630-648)
 *     and the node recursion of examples/gltfReader/main.cpp:1538-1576;
 *   - CadR::operator*(mat4, BoundingSphere) (src/CadR/BoundingSphere.h:70-87);
 *   - glm::lookAtLH (ext/matrix_transform.inl:122-142) and the
 *     perspectiveLH_ZO scale terms (ext/matrix_clip_space.inl:255-268);
 *   - the cull / LOD / per-StateSet compaction definition stated in
 *     include/cadr_hip.h (no reference implementation exists), with the
 *     StateSet draw ranges of src/CadR/StateSet.cpp:198-268.
 * Only the byte layouts are taken from include/cadr_hip.h.
 */
#ifndef CADR_ORACLE_H
#define CADR_ORACLE_H

#include <stdint.h>

#include "../include/cadr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_region {
	uint64_t base;   /* device address of the first byte */
	uint64_t size;
	uint8_t* host;   /* host copy */
} oracle_region;

/* Gather over the host image; returns 0, or -1 (and stops) when an address
 * the reference would dereference lies outside every region. */
int oracle_process_drawables(oracle_region* regions, uint32_t numRegions,
                             const cadrProcessDrawablesPushConstants* pc, uint32_t handleLevel,
                             uint64_t numDrawables);
/* lookupHandle of processDrawables.comp:77-89; -1 on an address outside the image. */
int oracle_lookup_handle(const oracle_region* regions, uint32_t numRegions, uint64_t root,
                         uint32_t handleLevel, uint64_t handle, uint64_t* address);

void oracle_mat4_mul(const float* A, const float* B, float* R);
void oracle_sphere_transform(const float* M, const float* sphere, float* out);
/* glTF node matrix from scale / rotation (NULL: none) / translation (NULL:
 * none) and the root-node axis flip (examples/gltfReader/main.cpp:1455-1553). */
void oracle_gltf_trs_matrix(const float* s, const float* q, const float* t, float* M);
void oracle_gltf_root_flip(const float* M, float* out);
/* glTF primitive and mesh bounding spheres (examples/gltfReader/main.cpp:
 * 2714, 3191-3245, 3597-3600, 3762-3769): out = {centre xyz, radius}.
 * positions: n x 3; boxes: k x {min xyz, max xyz}; spheres: k x 4. */
void oracle_gltf_primitive_sphere(const float* positions, uint64_t n, const float* boxMin, const float* boxMax,
                                  float* out);
void oracle_gltf_mesh_sphere(const float* boxes, const float* spheres, uint64_t k, float* out);
/* Threads used by oracle_propagate_level (default 1). */
void oracle_set_threads(int n);
/* One level: world = parentWorld[parent[i]] * local[i]; parentWorld == NULL
 * or parent[i] >= parentCount copies local.  spheres may be NULL. */
void oracle_propagate_level(const float* local, const uint32_t* parent, uint32_t parentCount,
                            const float* parentWorld, float* world, uint64_t count,
                            const uint32_t* meshIndex, const cadrMesh* meshes, uint32_t numMeshes,
                            float* spheres);
int oracle_make_view(const float eye[3], const float dir[3], const float up[3], float fovy, float aspect,
                     float zNear, float zFar, float viewportHeight, const float* lodThresholds,
                     cadrView* out);
/* Visibility and LOD of one drawable (lod < 0 when not visible). */
int oracle_test_drawable(const float* sphere, const cadrView* view, const cadrMesh* mesh);
/* Cull numViews views of one scene; commands[v*commandStride + ...], counts[v*K + k].
 * nthreads <= 1 runs single-threaded. */
void oracle_cull(const float* spheres, const uint32_t* meshIndex, const cadrMesh* meshes, uint32_t numMeshes,
                 const uint32_t* stateSetCounts, uint32_t numStateSets, const cadrView* views,
                 uint32_t numViews, uint32_t commandStride, cadrDrawIndexedIndirectCommand* commands,
                 uint32_t* counts, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
