// Synthetic scene levels generated in device memory (cadrSynthesizeLevel).
//
// Workload generation only: a pure function of (seed, stream, tree node) per
// node, so scenes of hundreds of millions of drawables appear in HBM without
// host arrays or PCIe traffic.  The arithmetic follows the host generator
// (cadr_amd/scenes.py) operation for operation, in double where it uses
// double, rounded to float at the same points.
#include <algorithm>

#include "internal.h"
#include "fp32.h"
#include "../../include/cadr_synth.h"

namespace cadr {
namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t randU64(uint64_t seed, uint32_t stream, uint64_t item, uint32_t comps, uint32_t comp)
{
	const uint64_t ctr = (uint64_t(stream) << 40) + item * comps + comp + 1u;
	return mix64(seed + ctr * kGamma);
}

__device__ __forceinline__ float randUnit(uint64_t seed, uint32_t stream, uint64_t item, uint32_t comps, uint32_t comp)
{
	return float(uint32_t(randU64(seed, stream, item, comps, comp) >> 40)) * (1.0f / 16777216.0f);
}

// 4-round Feistel network on the smallest even-bit domain covering n,
// cycle-walked into [0, n): a bijection of [0, n).
__host__ __device__ __forceinline__ uint64_t permute(uint64_t i, uint64_t n, uint64_t key)
{
	if (n <= 1) return i;
	uint32_t bits = 64u - uint32_t(__builtin_clzll(n - 1));
	bits = bits < 2u ? 2u : bits + (bits & 1u);
	const uint32_t half = bits / 2u;
	const uint64_t mask = (uint64_t(1) << half) - 1u;
	uint64_t x = i;
	do {
		for (uint32_t r = 0; r < 4u; r++) {
			const uint64_t L = x >> half, R = x & mask;
			const uint64_t F = mix64(key + (R + (uint64_t(r) << 40)) * kGamma) & mask;
			x = (R << half) | (L ^ F);
		}
	} while (x >= n);
	return x;
}

// Shoemake's uniform rotation from three uniforms, plus a translation
// (scenes.py _rotation_translation_chunk).
__device__ void rotationTranslation(uint64_t seed, uint32_t stream, uint64_t t, float range, float4 M[4])
{
	const double u1 = randUnit(seed, stream, t, 6, 0);
	const double u2 = randUnit(seed, stream, t, 6, 1);
	const double u3 = randUnit(seed, stream, t, 6, 2);
	const double a = sqrt(1.0 - u1), b = sqrt(u1);
	const double twoPi = 2.0 * 3.141592653589793;
	const double x = a * sin(twoPi * u2), y = a * cos(twoPi * u2);
	const double z = b * sin(twoPi * u3), w = b * cos(twoPi * u3);
	M[0] = make_float4(float(1.0 - 2.0 * (y * y + z * z)), float(2.0 * (x * y + z * w)), float(2.0 * (x * z - y * w)), 0.f);
	M[1] = make_float4(float(2.0 * (x * y - z * w)), float(1.0 - 2.0 * (x * x + z * z)), float(2.0 * (y * z + x * w)), 0.f);
	M[2] = make_float4(float(2.0 * (x * z + y * w)), float(2.0 * (y * z - x * w)), float(1.0 - 2.0 * (x * x + y * y)), 0.f);
	const float lo = -range, span = 2.0f * range;
	M[3] = make_float4(lo + span * randUnit(seed, stream, t, 6, 3), lo + span * randUnit(seed, stream, t, 6, 4),
	                   lo + span * randUnit(seed, stream, t, 6, 5), 1.f);
}

// Uniform scale exp(U[ln 0.1, ln 2]) plus a translation (scenes.py scale_translation).
__device__ void scaleTranslation(uint64_t seed, uint32_t stream, uint64_t t, float range, float4 M[4])
{
	const double lnLo = -2.3025850929940455;  // log(0.1), correctly rounded
	const double lnHi = 0.6931471805599453;   // log(2.0), correctly rounded
	const float s = float(exp(lnLo + double(randUnit(seed, stream, t, 4, 0)) * (lnHi - lnLo)));
	M[0] = make_float4(s, 0.f, 0.f, 0.f);
	M[1] = make_float4(0.f, s, 0.f, 0.f);
	M[2] = make_float4(0.f, 0.f, s, 0.f);
	const float lo = -range, span = 2.0f * range;
	M[3] = make_float4(lo + span * randUnit(seed, stream, t, 4, 1), lo + span * randUnit(seed, stream, t, 4, 2),
	                   lo + span * randUnit(seed, stream, t, 4, 3), 1.f);
}

__global__ void __launch_bounds__(256) synthLevelKernel(cadrSynthLevel L)
{
	const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
	for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < L.count; i += stride) {
		const uint64_t t = L.permuteKey ? permute(i, L.count, L.permuteKey) : i;
		uint32_t m = 0;
		if (L.meshIndex || L.spheres) {
			const uint64_t u = randU64(L.seed, L.meshStream, t, 1, 0) >> 32;
			m = uint32_t((u * uint64_t(L.numMeshes)) >> 32);
		}
		if (L.meshIndex) reinterpret_cast<uint32_t*>(L.meshIndex)[i] = m;
		if (L.parentIndex) reinterpret_cast<uint32_t*>(L.parentIndex)[i] = uint32_t(t / L.fanout);
		if (L.localMatrices || L.spheres) {
			float4 M[4];
			if (L.transform == CADR_SYNTH_ROTATION_TRANSLATION)
				rotationTranslation(L.seed, L.stream, t, L.translationRange, M);
			else
				scaleTranslation(L.seed, L.stream, t, L.translationRange, M);
			if (L.localMatrices) {
				float4* dst = reinterpret_cast<float4*>(L.localMatrices) + 4 * i;
#pragma unroll
				for (int c = 0; c < 4; c++) dst[c] = M[c];
			}
			if (L.spheres) {
				const float4 local = *reinterpret_cast<const float4*>(reinterpret_cast<const cadrMesh*>(L.meshes) + m);
				reinterpret_cast<float4*>(L.spheres)[i] = sphereTransform(M, local);
			}
		}
	}
}

}  // namespace
}  // namespace cadr

using namespace cadr;

CADR_API uint64_t cadrSynthPermute(uint64_t index, uint64_t count, uint64_t key)
{
	return key ? permute(index, count, key) : index;
}

CADR_API int cadrSynthesizeLevel(const cadrSynthLevel* level, cadrStream stream)
{
	CADR_REQUIRE(level, "cadrSynthesizeLevel: level is NULL");
	const cadrSynthLevel L = *level;
	if (L.count == 0) return CADR_SUCCESS;
	CADR_REQUIRE(L.count < (uint64_t(1) << 40), "cadrSynthesizeLevel: count too large");
	CADR_REQUIRE(L.transform <= CADR_SYNTH_SCALE_TRANSLATION, "cadrSynthesizeLevel: unknown transform kind");
	CADR_REQUIRE(L.localMatrices % 16 == 0 && L.spheres % 16 == 0 && L.meshes % 16 == 0 && L.parentIndex % 4 == 0 &&
	                 L.meshIndex % 4 == 0,
	             "cadrSynthesizeLevel: misaligned output");
	CADR_REQUIRE(!L.parentIndex || L.fanout > 0, "cadrSynthesizeLevel: parentIndex needs fanout > 0");
	CADR_REQUIRE(!(L.meshIndex || L.spheres) || L.numMeshes > 0, "cadrSynthesizeLevel: mesh outputs need numMeshes > 0");
	CADR_REQUIRE(!L.spheres || L.meshes, "cadrSynthesizeLevel: spheres need the mesh table");
	const uint64_t blocks = std::min<uint64_t>((L.count + 255) / 256, uint64_t(computeUnitCount()) * 16u);
	hipLaunchKernelGGL(synthLevelKernel, dim3(uint32_t(blocks)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), L);
	CADR_CHECK_HIP(hipGetLastError(), "synthLevelKernel launch");
	return CADR_SUCCESS;
}
