// Host-side view construction: camera -> cadrView constants.
//
// The view matrix is glm::lookAtLH (3rdParty/GLM/glm/ext/matrix_transform.inl:122-142)
// evaluated in float with glm's own operation order:
//   normalize(v) = v * (1 / sqrt(dot(v, v)))   (func_geometric.inl, func_exponential.inl)
//   dot(a, b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(x, y)  = (x.y*y.z - y.y*x.z, x.z*y.x - y.z*x.x, x.x*y.y - y.x*x.y)
// and the projection scale P[1][1] = 1/tan(fovy/2), P[0][0] = 1/(aspect*tan(fovy/2))
// is glm::perspectiveLH_ZO (ext/matrix_clip_space.inl:255-268).  The side
// planes of the frustum through the eye have unit normals built from those
// two entries: {cos, sin} = {P/sqrt(P*P + 1), 1/sqrt(P*P + 1)}.
//
// This translation unit is compiled with -ffp-contract=off like the kernels.
#include "internal.h"

#include <cmath>

namespace {

struct V3 {
	float x, y, z;
};

inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 x, V3 y)
{
	return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline V3 normalize(V3 v)
{
	const float inv = 1.0f / std::sqrt(dot(v, v));
	return {v.x * inv, v.y * inv, v.z * inv};
}

}  // namespace

CADR_API int cadrMakeView(const float eye[3], const float dir[3], const float up[3], float fovy, float aspect,
                          float zNear, float zFar, float viewportHeight, const float* lodThresholds, cadrView* out)
{
	CADR_REQUIRE(eye && dir && up && out, "cadrMakeView: null argument");
	CADR_REQUIRE(fovy > 0.0f && fovy < 3.14159265f, "cadrMakeView: fovy must be in (0, pi)");
	CADR_REQUIRE(aspect > 0.0f && zNear >= 0.0f && zFar > zNear && viewportHeight > 0.0f,
	             "cadrMakeView: invalid aspect, depth range or viewport height");
	const V3 e{eye[0], eye[1], eye[2]};
	const V3 d{dir[0], dir[1], dir[2]};
	const V3 u0{up[0], up[1], up[2]};
	CADR_REQUIRE(dot(d, d) > 0.0f, "cadrMakeView: zero view direction");

	// glm::lookAtLH(eye, center = eye + dir, up)
	const V3 center = add(e, d);
	const V3 f = normalize(sub(center, e));
	const V3 s = normalize(cross(u0, f));
	const V3 u = cross(f, s);
	CADR_REQUIRE(std::isfinite(s.x) && std::isfinite(s.y) && std::isfinite(s.z),
	             "cadrMakeView: up is parallel to the view direction");

	cadrView v{};
	// rows of the world->view matrix: {V[0][r], V[1][r], V[2][r], V[3][r]}
	v.viewRow[0][0] = s.x; v.viewRow[0][1] = s.y; v.viewRow[0][2] = s.z; v.viewRow[0][3] = -dot(s, e);
	v.viewRow[1][0] = u.x; v.viewRow[1][1] = u.y; v.viewRow[1][2] = u.z; v.viewRow[1][3] = -dot(u, e);
	v.viewRow[2][0] = f.x; v.viewRow[2][1] = f.y; v.viewRow[2][2] = f.z; v.viewRow[2][3] = -dot(f, e);

	// glm::perspectiveLH_ZO scale terms
	const float tanHalfFovy = std::tan(fovy / 2.0f);
	const float p00 = 1.0f / (aspect * tanHalfFovy);
	const float p11 = 1.0f / tanHalfFovy;
	const float nx = std::sqrt(p00 * p00 + 1.0f);
	const float ny = std::sqrt(p11 * p11 + 1.0f);
	v.nearZ = zNear;
	v.farZ = zFar;
	v.sideX[0] = p00 / nx;
	v.sideX[1] = 1.0f / nx;
	v.sideY[0] = p11 / ny;
	v.sideY[1] = 1.0f / ny;
	v.lodScale = (0.5f * viewportHeight) * p11;
	if (lodThresholds) {
		v.lodThreshold[0] = lodThresholds[0];
		v.lodThreshold[1] = lodThresholds[1];
		v.lodThreshold[2] = lodThresholds[2];
	} else {
		v.lodThreshold[0] = 64.0f;
		v.lodThreshold[1] = 8.0f;
		v.lodThreshold[2] = 1.0f;
	}
	*out = v;
	return CADR_SUCCESS;
}

static_assert(sizeof(cadrView) == 96, "cadrView is 96 bytes");
static_assert(sizeof(cadrMesh) == 64, "cadrMesh is 64 bytes");
static_assert(sizeof(cadrCullJob) == 80, "cadrCullJob is 80 bytes");
static_assert(sizeof(cadrCullTile) == 32, "cadrCullTile is 32 bytes");
static_assert(sizeof(cadrDrawIndexedIndirectCommand) == 20, "VkDrawIndexedIndirectCommand is 20 bytes");
static_assert(sizeof(cadrTransformLevel) == 72, "cadrTransformLevel is 72 bytes");
static_assert(sizeof(cadrDrawableGpuData) == 48, "DrawableGpuData is 48 bytes");
