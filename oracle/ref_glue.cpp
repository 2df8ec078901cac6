// C entry points around the reference's own FP32 math, compiled from the
// reference tree where it lies (src/CadR/BoundingSphere.h, 3rdParty/GLM) by
// oracle/Makefile into oracle/_ref/libcadr_ref.so.  TEST INFRASTRUCTURE ONLY:
// it pins the oracle's restatement of these functions bit for bit.
#include <CadR/BoundingSphere.h>
#include <glm/ext/matrix_clip_space.hpp>
#include <glm/ext/matrix_transform.hpp>
#include <glm/gtc/quaternion.hpp>
#include <glm/mat3x3.hpp>
#include <glm/mat4x4.hpp>

#include <cstdint>
#include <cstring>

namespace {

glm::mat4 loadMat(const float* p)
{
	glm::mat4 m;
	std::memcpy(&m[0][0], p, 64);
	return m;
}

void storeMat(const glm::mat4& m, float* p) { std::memcpy(p, &m[0][0], 64); }

}  // namespace

extern "C" {

// glm operator*(mat4, mat4), type_mat4x4.inl:630-648
void ref_mat4_mul(const float* A, const float* B, float* R) { storeMat(loadMat(A) * loadMat(B), R); }

// glm operator*(mat4, vec4), type_mat4x4.inl:536-572
void ref_mat4_vec4(const float* M, const float* v, float* out)
{
	const glm::vec4 r = loadMat(M) * glm::vec4(v[0], v[1], v[2], v[3]);
	out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

// CadR::operator*(mat4, BoundingSphere), BoundingSphere.h:70-87
void ref_sphere_transform(const float* M, const float* s, float* out)
{
	const CadR::BoundingSphere b{glm::vec3(s[0], s[1], s[2]), s[3]};
	const CadR::BoundingSphere r = loadMat(M) * b;
	out[0] = r.center.x; out[1] = r.center.y; out[2] = r.center.z; out[3] = r.radius;
}

// glm::lookAtLH, ext/matrix_transform.inl:122-142
void ref_look_at_lh(const float* eye, const float* center, const float* up, float* out)
{
	storeMat(glm::lookAtLH(glm::vec3(eye[0], eye[1], eye[2]), glm::vec3(center[0], center[1], center[2]),
	                       glm::vec3(up[0], up[1], up[2])),
	         out);
}

// glm::perspectiveLH_ZO, ext/matrix_clip_space.inl:255-268
void ref_perspective_lh_zo(float fovy, float aspect, float zNear, float zFar, float* out)
{
	storeMat(glm::perspectiveLH_ZO(fovy, aspect, zNear, zFar), out);
}

// Depth-first node propagation in the manner of examples/gltfReader/main.cpp:1538-1576
// (m = parentMatrix * node.matrix, glm operator*) for a tree given as
// parent indices in topological order (parent[i] < i, or -1 for a root):
// each node's world matrix is its parent's world matrix times its local one.
void ref_propagate_tree(const float* local, const int64_t* parent, uint64_t count, float* world)
{
	for (uint64_t i = 0; i < count; i++) {
		const glm::mat4 m = parent[i] < 0 ? loadMat(local + 16 * i)
		                                  : loadMat(world + 16 * parent[i]) * loadMat(local + 16 * i);
		storeMat(m, world + 16 * i);
	}
}

// glTF node matrix as examples/gltfReader/main.cpp:1455-1500 builds it from
// scale, rotation (glm::mat3 of a glm::quat) and translation; q or t may be
// null (absent in the node).
void ref_gltf_trs_matrix(const float* s, const float* q, const float* t, float* out)
{
	const glm::vec3 scale(s[0], s[1], s[2]);
	glm::mat4 m(0.f);
	if (q) {
		glm::quat quat;
		quat.x = q[0];
		quat.y = q[1];
		quat.z = q[2];
		quat.w = q[3];
		const glm::mat3 r = glm::mat3(quat);
		m[0] = glm::vec4(r[0] * scale.x, 0.f);
		m[1] = glm::vec4(r[1] * scale.y, 0.f);
		m[2] = glm::vec4(r[2] * scale.z, 0.f);
		m[3] = glm::vec4(0.f, 0.f, 0.f, 1.f);
	} else {
		m[0][0] = scale.x;
		m[1][1] = scale.y;
		m[2][2] = scale.z;
		m[3][3] = 1.f;
	}
	if (t) {
		m[3][0] = t[0];
		m[3][1] = t[1];
		m[3][2] = t[2];
	}
	storeMat(m, out);
}

// Root matrix of a glTF scene (main.cpp:1547-1553): y and z rows negated.
// glTF primitive sphere as examples/gltfReader/main.cpp:3196-3245, 3597-3600
// builds it with the reference's BoundingBox / BoundingSphere (main.cpp has
// `using namespace std`, so its sqrt of a float is std::sqrt(float)).
void ref_gltf_primitive_sphere(const float* positions, uint64_t n, const float* boxMin, const float* boxMax,
                               float* out)
{
	CadR::BoundingBox bb{glm::vec3(boxMin[0], boxMin[1], boxMin[2]), glm::vec3(boxMax[0], boxMax[1], boxMax[2])};
	CadR::BoundingSphere bs{bb.getCenter(), 0.f};
	for (uint64_t i = 0; i < n; i++)
		bs.extendRadiusByPointUsingRadius2(glm::vec3(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]));
	bs.radius = std::sqrt(bs.radius);
	out[0] = bs.center.x; out[1] = bs.center.y; out[2] = bs.center.z; out[3] = bs.radius;
}

// glTF mesh sphere as main.cpp:2714, 3195, 3762-3769 builds it.
void ref_gltf_mesh_sphere(const float* boxes, const float* spheres, uint64_t k, float* out)
{
	CadR::BoundingBox bb = CadR::BoundingBox::empty();
	for (uint64_t i = 0; i < k; i++)
		bb.extendBy(CadR::BoundingBox{glm::vec3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]),
		                              glm::vec3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5])});
	CadR::BoundingSphere bs{bb.getCenter(), 0.f};
	for (uint64_t i = 0; i < k; i++)
		bs.extendRadiusBy(CadR::BoundingSphere{glm::vec3(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2]),
		                                       spheres[4 * i + 3]});
	out[0] = bs.center.x; out[1] = bs.center.y; out[2] = bs.center.z; out[3] = bs.radius;
}

void ref_gltf_root_flip(const float* M, float* out)
{
	glm::mat4 m = loadMat(M);
	for (int i = 0; i < 4; i++) {
		m[i][1] = -m[i][1];
		m[i][2] = -m[i][2];
	}
	storeMat(m, out);
}

}  // extern "C"
