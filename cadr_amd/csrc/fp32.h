// FP32 helpers with the evaluation order of the reference's glm 0.9.9 math.
//
// Every file including this header is compiled with -ffp-contract=off and
// without fast-math, so a*b + c*d is two roundings of products and one of the
// sum, exactly as the reference's x86-64 build (ISO C++17, no FMA) computes it.
#pragma once

#include <hip/hip_runtime.h>

namespace cadr {

// glm mat4 * mat4, 3rdParty/GLM/glm/detail/type_mat4x4.inl:630-648:
//   R[j] = ((A[0]*B[j][0] + A[1]*B[j][1]) + A[2]*B[j][2]) + A[3]*B[j][3]
__device__ __forceinline__ void mat4Mul(const float4 A[4], const float4 B[4], float4 R[4])
{
#pragma unroll
	for (int j = 0; j < 4; j++) {
		const float4 b = B[j];
		float4 r;
		r.x = ((A[0].x * b.x + A[1].x * b.y) + A[2].x * b.z) + A[3].x * b.w;
		r.y = ((A[0].y * b.x + A[1].y * b.y) + A[2].y * b.z) + A[3].y * b.w;
		r.z = ((A[0].z * b.x + A[1].z * b.y) + A[2].z * b.z) + A[3].z * b.w;
		r.w = ((A[0].w * b.x + A[1].w * b.y) + A[2].w * b.z) + A[3].w * b.w;
		R[j] = r;
	}
}

// std::max as BoundingSphere.h uses it: (a < b) ? b : a.
__device__ __forceinline__ float stdMax(float a, float b) { return (a < b) ? b : a; }

// World bounding sphere, src/CadR/BoundingSphere.h:70-87:
//   c' = mat3(M) * c + vec3(M[3])    (mat3*vec3: m[0][i]*x + m[1][i]*y + m[2][i]*z, left to right)
//   r' = sqrt(max(max(|M0|^2, |M1|^2), |M2|^2)) * r,  |v|^2 = x*x + y*y + z*z
// The squared lengths and their maximum are float.  The reference's
// unqualified sqrt(float) inside namespace CadR resolves to ::sqrt(double), so
// the square root and the product with r are double and the result is
// rounded to float once: r' = float(sqrt(double(max)) * double(r)).  The
// device f64 sqrt expansion is correctly rounded, like the host's.
__device__ __forceinline__ float4 sphereTransform(const float4 M[4], float4 s)
{
	float4 o;
	o.x = ((M[0].x * s.x + M[1].x * s.y) + M[2].x * s.z) + M[3].x;
	o.y = ((M[0].y * s.x + M[1].y * s.y) + M[2].y * s.z) + M[3].y;
	o.z = ((M[0].z * s.x + M[1].z * s.y) + M[2].z * s.z) + M[3].z;
	const float l0 = (M[0].x * M[0].x + M[0].y * M[0].y) + M[0].z * M[0].z;
	const float l1 = (M[1].x * M[1].x + M[1].y * M[1].y) + M[1].z * M[1].z;
	const float l2 = (M[2].x * M[2].x + M[2].y * M[2].y) + M[2].z * M[2].z;
	o.w = float(sqrt(double(stdMax(stdMax(l0, l1), l2))) * double(s.w));
	return o;
}

}  // namespace cadr
