/*
 * CPU oracle of the drawable-processing pass — TEST INFRASTRUCTURE ONLY.
 * See cadr_oracle.h for scope; every function cites the reference lines it
 * restates.  Build: oracle/Makefile (gcc -std=c11 -O2 -ffp-contract=off).
 */
#include "cadr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- memory */

static uint8_t* resolve(const oracle_region* r, uint32_t n, uint64_t addr, uint64_t bytes)
{
	for (uint32_t i = 0; i < n; i++) {
		if (addr >= r[i].base && addr - r[i].base <= r[i].size && bytes <= r[i].size - (addr - r[i].base))
			return r[i].host + (addr - r[i].base);
	}
	return NULL;
}

static int load_u64(const oracle_region* r, uint32_t n, uint64_t addr, uint64_t* v)
{
	uint8_t* p = resolve(r, n, addr, 8);
	if (!p) return -1;
	memcpy(v, p, 8);
	return 0;
}

static int load_u32(const oracle_region* r, uint32_t n, uint64_t addr, uint32_t* v)
{
	uint8_t* p = resolve(r, n, addr, 4);
	if (!p) return -1;
	memcpy(v, p, 4);
	return 0;
}

/* processDrawables.comp:77-89.  HandleTableRef(x).pointers[i] reads the
 * uint64 at x + 8*i; uint(handle) truncates to 32 bits.  The product
 * resolves an index past the 2048-entry table, or a null table, to address 0
 * (the GLSL reads whatever lies there); the oracle models the same rule so
 * fixtures may contain such handles. */
int oracle_lookup_handle(const oracle_region* r, uint32_t n, uint64_t root, uint32_t level, uint64_t handle,
                         uint64_t* address)
{
	uint32_t h = (uint32_t)handle;
	uint64_t t2, t3;
	*address = 0;
	if (level == 1) {
		if (h >= CADR_HANDLES_PER_TABLE) return 0;
		return load_u64(r, n, root + 8ull * h, address);
	}
	if (level == 2) {
		uint32_t i1 = (uint32_t)(handle >> 11);
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		if (load_u64(r, n, root + 8ull * i1, &t2)) return -1;
		if (t2 == 0) return 0;
		return load_u64(r, n, t2 + 8ull * (h & 0x7ffu), address);
	}
	{
		uint32_t i1 = (uint32_t)(handle >> 22);
		if (i1 >= CADR_HANDLES_PER_TABLE) return 0;
		if (load_u64(r, n, root + 8ull * i1, &t2)) return -1;
		if (t2 == 0) return 0;
		if (load_u64(r, n, t2 + 8ull * ((uint32_t)(handle >> 11) & 0x7ffu), &t3)) return -1;
		if (t3 == 0) return 0;
		return load_u64(r, n, t3 + 8ull * (h & 0x7ffu), address);
	}
}

/* processDrawables.comp:92-113, one invocation per drawable i (the 32768 x Y
 * split of Renderer.cpp:684-692 enumerates i = y*32768 + x = 0..N-1). */
int oracle_process_drawables(oracle_region* r, uint32_t n, const cadrProcessDrawablesPushConstants* pc,
                             uint32_t level, uint64_t numDrawables)
{
	for (uint64_t i = 0; i < numDrawables; i++) {
		const uint64_t d = pc->drawableList + i * 48u;
		uint64_t vH, iH, mlH, ddH, psH;
		uint32_t psOffset;
		if (load_u64(r, n, d + 0, &vH) || load_u64(r, n, d + 8, &iH) || load_u64(r, n, d + 16, &mlH) ||
		    load_u64(r, n, d + 24, &ddH) || load_u64(r, n, d + 32, &psH) || load_u32(r, n, d + 40, &psOffset))
			return -1;
		uint64_t psBase, ml, vp, ip, dp;
		if (oracle_lookup_handle(r, n, pc->handleTable, level, psH, &psBase) ||
		    oracle_lookup_handle(r, n, pc->handleTable, level, mlH, &ml) ||
		    oracle_lookup_handle(r, n, pc->handleTable, level, vH, &vp) ||
		    oracle_lookup_handle(r, n, pc->handleTable, level, iH, &ip) ||
		    oracle_lookup_handle(r, n, pc->handleTable, level, ddH, &dp))
			return -1;
		uint32_t count = 0, first = 0, numMatrices = 0;
		if (psBase) {
			if (load_u32(r, n, psBase + psOffset, &count) || load_u32(r, n, psBase + psOffset + 4, &first))
				return -1;
		}
		if (ml) {
			if (load_u32(r, n, ml, &numMatrices)) return -1;
		}
		uint8_t* ind = resolve(r, n, pc->indirectData + i * 16u, 16);
		uint8_t* ptr = resolve(r, n, pc->drawablePointers + i * 32u, 32);
		if (!ind || !ptr) return -1;
		const uint32_t cmd[4] = {count, numMatrices, first, 0u};
		const uint64_t pts[4] = {vp, ip, ml, dp};
		memcpy(ind, cmd, 16);
		memcpy(ptr, pts, 32);
	}
	return 0;
}

/* ------------------------------------------------------------- FP32 math */

/* glm operator*(mat4, mat4), type_mat4x4.inl:643-646, column-major:
 * R[j] = SrcA0*SrcB[j][0] + SrcA1*SrcB[j][1] + SrcA2*SrcB[j][2] + SrcA3*SrcB[j][3],
 * each vec4 sum evaluated left to right. */
void oracle_mat4_mul(const float* A, const float* B, float* R)
{
	float out[16];
	for (int j = 0; j < 4; j++) {
		for (int r = 0; r < 4; r++) {
			float acc = A[0 * 4 + r] * B[j * 4 + 0];
			acc = acc + A[1 * 4 + r] * B[j * 4 + 1];
			acc = acc + A[2 * 4 + r] * B[j * 4 + 2];
			acc = acc + A[3 * 4 + r] * B[j * 4 + 3];
			out[j * 4 + r] = acc;
		}
	}
	memcpy(R, out, sizeof(out));
}

static float std_max(float a, float b) { return (a < b) ? b : a; } /* std::max */

/* BoundingSphere.h:70-87: center = glm::mat3(M) * c + glm::vec3(M[3]) with
 * mat3*vec3 = m[0][i]*x + m[1][i]*y + m[2][i]*z (type_mat3x3.inl:470-473);
 * radius = sqrt(max(max(sqrLength(M[0]), sqrLength(M[1])), sqrLength(M[2]))) * r
 * where the unqualified sqrt of BoundingSphere.h:84 is ::sqrt(double): the
 * root and the product are double, rounded to float by the float-returning
 * lambda. */
/* glTF node matrix from scale s, rotation quaternion q {x,y,z,w} (NULL: none)
 * and translation t (NULL: none), examples/gltfReader/main.cpp:1455-1500:
 * with a rotation, the columns of glm::mat3(q) (mat3_cast,
 * gtc/quaternion.inl:41-66) scaled by s.x, s.y, s.z, then w row 0,0,0,1; without
 * one, diag(s.x, s.y, s.z, 1); the translation, when given, replaces column 3's
 * x, y, z.  Column-major. */
void oracle_gltf_trs_matrix(const float* s, const float* q, const float* t, float* M)
{
	memset(M, 0, 64);
	if (q) {
		const float qxx = q[0] * q[0], qyy = q[1] * q[1], qzz = q[2] * q[2];
		const float qxz = q[0] * q[2], qxy = q[0] * q[1], qyz = q[1] * q[2];
		const float qwx = q[3] * q[0], qwy = q[3] * q[1], qwz = q[3] * q[2];
		float r[9];
		r[0] = 1.f - 2.f * (qyy + qzz);
		r[1] = 2.f * (qxy + qwz);
		r[2] = 2.f * (qxz - qwy);
		r[3] = 2.f * (qxy - qwz);
		r[4] = 1.f - 2.f * (qxx + qzz);
		r[5] = 2.f * (qyz + qwx);
		r[6] = 2.f * (qxz + qwy);
		r[7] = 2.f * (qyz - qwx);
		r[8] = 1.f - 2.f * (qxx + qyy);
		for (int c = 0; c < 3; c++)
			for (int k = 0; k < 3; k++) M[4 * c + k] = r[3 * c + k] * s[c];
	} else {
		M[0] = s[0];
		M[5] = s[1];
		M[10] = s[2];
	}
	M[15] = 1.f;
	if (t) {
		M[12] = t[0];
		M[13] = t[1];
		M[14] = t[2];
	}
}

/* Root matrix of a glTF scene, main.cpp:1547-1553: rows 1 and 2 (y and z)
 * negated in every column. */
void oracle_gltf_root_flip(const float* M, float* out)
{
	for (int c = 0; c < 4; c++) {
		out[4 * c + 0] = M[4 * c + 0];
		out[4 * c + 1] = -M[4 * c + 1];
		out[4 * c + 2] = -M[4 * c + 2];
		out[4 * c + 3] = M[4 * c + 3];
	}
}

/* glm::distance2(p0, p1) = length2(p1 - p0) = (d.x*d.x + d.y*d.y) + d.z*d.z
 * (gtx/norm.inl:9-37, detail/func_geometric.inl dot of vec3). */
static float distance2(const float* p0, const float* p1)
{
	const float dx = p1[0] - p0[0], dy = p1[1] - p0[1], dz = p1[2] - p0[2];
	return (dx * dx + dy * dy) + dz * dz;
}

/* Primitive sphere, main.cpp:3196-3245 and :3597-3600: centre = the POSITION
 * accessor box's getCenter() = (min + max) / 2 (BoundingBox.h:44); radius
 * holds radius^2 while each position p is added by
 * extendRadiusByPointUsingRadius2 (BoundingSphere.h:65-69: d2 =
 * distance2(p, centre), kept when d2 > radius, so NaN never is), then
 * radius = sqrt(radius) with main.cpp's `using namespace std`: the float
 * overload. */
void oracle_gltf_primitive_sphere(const float* positions, uint64_t n, const float* boxMin, const float* boxMax,
                                  float* out)
{
	float c[3];
	for (int a = 0; a < 3; a++) c[a] = (boxMin[a] + boxMax[a]) / 2.0f;
	float r2 = 0.0f;
	for (uint64_t i = 0; i < n; i++) {
		const float d2 = distance2(positions + 3 * i, c);
		if (d2 > r2) r2 = d2;
	}
	out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
	out[3] = sqrtf(r2);
}

/* Mesh sphere, main.cpp:2714, :3195, :3762-3769: the union of the
 * primitives' boxes (BoundingBox::empty() then extendBy, BoundingBox.h:
 * 30-35, 55-68: a component is replaced only by a smaller min / larger max),
 * centred there with radius 0 and grown by each primitive sphere in order
 * (extendRadiusBy, BoundingSphere.h:50-58: d2 = distance2(centre, s.centre),
 * dif = radius - s.radius, when dif < 0 or d2 > dif*dif the radius becomes
 * sqrt(d2) + s.radius, that unqualified sqrt inside namespace CadR being
 * ::sqrt(double), so the sum is rounded to float once). */
void oracle_gltf_mesh_sphere(const float* boxes, const float* spheres, uint64_t k, float* out)
{
	float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
	for (uint64_t i = 0; i < k; i++)
		for (int a = 0; a < 3; a++) {
			if (boxes[6 * i + a] < mn[a]) mn[a] = boxes[6 * i + a];
			if (boxes[6 * i + 3 + a] > mx[a]) mx[a] = boxes[6 * i + 3 + a];
		}
	float c[3];
	for (int a = 0; a < 3; a++) c[a] = (mn[a] + mx[a]) / 2.0f;
	float r = 0.0f;
	for (uint64_t i = 0; i < k; i++) {
		const float d2 = distance2(c, spheres + 4 * i);
		const float dif = r - spheres[4 * i + 3];
		if (dif < 0.0f || d2 > dif * dif) r = (float)(sqrt((double)d2) + (double)spheres[4 * i + 3]);
	}
	out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
	out[3] = r;
}

void oracle_sphere_transform(const float* M, const float* s, float* out)
{
	float c[3];
	for (int i = 0; i < 3; i++) {
		float t = M[0 * 4 + i] * s[0];
		t = t + M[1 * 4 + i] * s[1];
		t = t + M[2 * 4 + i] * s[2];
		c[i] = t + M[3 * 4 + i];
	}
	float l[3];
	for (int k = 0; k < 3; k++) {
		const float x = M[k * 4 + 0], y = M[k * 4 + 1], z = M[k * 4 + 2];
		l[k] = x * x + y * y + z * z;
	}
	const float maxSqr = std_max(std_max(l[0], l[1]), l[2]);
	out[0] = c[0];
	out[1] = c[1];
	out[2] = c[2];
	out[3] = (float)(sqrt((double)maxSqr) * (double)s[3]);
}

/* examples/gltfReader/main.cpp:1538-1576 (m = parentMatrix * node.matrix),
 * one tree level at a time. */
/* Threads of oracle_propagate_level (1 by default); nodes are independent, so
 * the results do not depend on it. */
static int g_propagate_threads = 1;

void oracle_set_threads(int n) { g_propagate_threads = n < 1 ? 1 : n; }

void oracle_propagate_level(const float* local, const uint32_t* parent, uint32_t parentCount,
                            const float* parentWorld, float* world, uint64_t count, const uint32_t* meshIndex,
                            const cadrMesh* meshes, uint32_t numMeshes, float* spheres)
{
#pragma omp parallel for schedule(static) num_threads(g_propagate_threads) if (g_propagate_threads > 1)
	for (uint64_t i = 0; i < count; i++) {
		float* W = world + 16 * i;
		if (parentWorld && parent[i] < parentCount)
			oracle_mat4_mul(parentWorld + 16ull * parent[i], local + 16 * i, W);
		else
			memcpy(W, local + 16 * i, 64);
		if (spheres) {
			const uint32_t m = meshIndex[i];
			if (m >= numMeshes) {
				/* BoundingSphere::empty(): center 0, radius -inf */
				spheres[4 * i + 0] = 0.f;
				spheres[4 * i + 1] = 0.f;
				spheres[4 * i + 2] = 0.f;
				spheres[4 * i + 3] = -INFINITY;
			} else {
				const float ls[4] = {meshes[m].center[0], meshes[m].center[1], meshes[m].center[2],
				                     meshes[m].radius};
				oracle_sphere_transform(W, ls, spheres + 4 * i);
			}
		}
	}
}

/* glm::lookAtLH (ext/matrix_transform.inl:122-142) with glm's dot, cross and
 * normalize (func_geometric.inl), and perspectiveLH_ZO's P[0][0], P[1][1]
 * (ext/matrix_clip_space.inl:259-266); frustum side constants and LOD scale
 * as defined in include/cadr_hip.h. */
static float dot3(const float* a, const float* b)
{
	float t = a[0] * b[0];
	t = t + a[1] * b[1];
	return t + a[2] * b[2];
}

static void cross3(const float* x, const float* y, float* o)
{
	o[0] = x[1] * y[2] - y[1] * x[2];
	o[1] = x[2] * y[0] - y[2] * x[0];
	o[2] = x[0] * y[1] - y[0] * x[1];
}

static void normalize3(const float* v, float* o)
{
	const float inv = 1.0f / sqrtf(dot3(v, v));
	o[0] = v[0] * inv;
	o[1] = v[1] * inv;
	o[2] = v[2] * inv;
}

int oracle_make_view(const float eye[3], const float dir[3], const float up[3], float fovy, float aspect,
                     float zNear, float zFar, float viewportHeight, const float* th, cadrView* out)
{
	float center[3], fc[3], f[3], su[3], s[3], u[3];
	for (int i = 0; i < 3; i++) center[i] = eye[i] + dir[i];
	for (int i = 0; i < 3; i++) fc[i] = center[i] - eye[i];
	normalize3(fc, f);
	cross3(up, f, su);
	normalize3(su, s);
	cross3(f, s, u);
	memset(out, 0, sizeof(*out));
	for (int i = 0; i < 3; i++) {
		out->viewRow[0][i] = s[i];
		out->viewRow[1][i] = u[i];
		out->viewRow[2][i] = f[i];
	}
	out->viewRow[0][3] = -dot3(s, eye);
	out->viewRow[1][3] = -dot3(u, eye);
	out->viewRow[2][3] = -dot3(f, eye);
	const float t = tanf(fovy / 2.0f);
	const float p00 = 1.0f / (aspect * t);
	const float p11 = 1.0f / t;
	const float nx = sqrtf(p00 * p00 + 1.0f);
	const float ny = sqrtf(p11 * p11 + 1.0f);
	out->nearZ = zNear;
	out->farZ = zFar;
	out->sideX[0] = p00 / nx;
	out->sideX[1] = 1.0f / nx;
	out->sideY[0] = p11 / ny;
	out->sideY[1] = 1.0f / ny;
	out->lodScale = (0.5f * viewportHeight) * p11;
	out->lodThreshold[0] = th ? th[0] : 64.0f;
	out->lodThreshold[1] = th ? th[1] : 8.0f;
	out->lodThreshold[2] = th ? th[2] : 1.0f;
	return 0;
}

/* ------------------------------------------------------------------ cull */

/* View-space coordinate with glm mat4*vec4 order (type_mat4x4.inl:561-572):
 * (m0*x + m1*y) + (m2*z + m3*1), the last product exact for w = 1. */
static float view_coord(const float* row, const float* c)
{
	const float a = row[0] * c[0];
	const float b = row[1] * c[1];
	const float d = row[2] * c[2];
	const float e = row[3];
	return (a + b) + (d + e);
}

int oracle_test_drawable(const float* s, const cadrView* v, const cadrMesh* mesh)
{
	const float vx = view_coord(v->viewRow[0], s);
	const float vy = view_coord(v->viewRow[1], s);
	const float vz = view_coord(v->viewRow[2], s);
	const float r = s[3];
	if (!(vz + r >= v->nearZ)) return -1;
	if (!(vz - r <= v->farZ)) return -1;
	if (!(fabsf(vx) * v->sideX[0] - vz * v->sideX[1] <= r)) return -1;
	if (!(fabsf(vy) * v->sideY[0] - vz * v->sideY[1] <= r)) return -1;
	const float p = r * v->lodScale;
	int lod = 0;
	for (int k = 0; k < 3; k++)
		if (p < v->lodThreshold[k] * vz) lod++;
	uint32_t lc = mesh->lodCount;
	if (lc < 1) lc = 1;
	if (lc > 4) lc = 4;
	if ((uint32_t)lod > lc - 1) lod = (int)(lc - 1);
	return lod;
}

void oracle_cull(const float* spheres, const uint32_t* meshIndex, const cadrMesh* meshes, uint32_t numMeshes,
                 const uint32_t* stateSetCounts, uint32_t K, const cadrView* views, uint32_t numViews,
                 uint32_t commandStride, cadrDrawIndexedIndirectCommand* commands, uint32_t* counts, int nthreads)
{
	uint64_t* base = (uint64_t*)malloc(sizeof(uint64_t) * (K + 1));
	base[0] = 0;
	for (uint32_t k = 0; k < K; k++) base[k + 1] = base[k] + stateSetCounts[k];
	const int64_t pairs = (int64_t)numViews * K;
#ifdef _OPENMP
	if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
	for (int64_t q = 0; q < pairs; q++) {
		const uint32_t v = (uint32_t)(q / K), k = (uint32_t)(q % K);
		const cadrView* view = views + v;
		cadrDrawIndexedIndirectCommand* out = commands + (uint64_t)v * commandStride;
		uint32_t c = 0;
		for (uint64_t i = base[k]; i < base[k + 1]; i++) {
			const uint32_t m = meshIndex[i];
			if (m >= numMeshes) continue;
			const int lod = oracle_test_drawable(spheres + 4 * i, view, meshes + m);
			if (lod < 0) continue;
			cadrDrawIndexedIndirectCommand* o = out + base[k] + c;
			o->indexCount = meshes[m].indexCount[lod];
			o->instanceCount = 1;
			o->firstIndex = meshes[m].firstIndex[lod];
			o->vertexOffset = meshes[m].vertexOffset;
			o->firstInstance = (uint32_t)i;
			c++;
		}
		counts[(uint64_t)v * K + k] = c;
	}
	free(base);
}
