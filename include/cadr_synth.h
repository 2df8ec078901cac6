/* Synthetic scenes generated directly in device memory.
 *
 * This header has no counterpart in the reference: it builds the benchmark
 * and test workloads (up to hundreds of millions of drawables) in HBM without
 * host arrays of that size.  Every value is a pure function of (seed, stream,
 * item, component) through a counter-based SplitMix64:
 *
 *   ctr = (stream << 40) + item * comps + comp + 1
 *   x   = seed + ctr * 0x9E3779B97F4A7C15
 *   z   = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9
 *   z   = (z ^ (z >> 27)) * 0x94D049BB133111EB
 *   u64 = z ^ (z >> 31)
 *
 * uniform float in [0,1) = float(u64 >> 40) * 2^-24; uniform integer in
 * [lo, hi] = lo + (((u64 >> 32) * (hi - lo + 1)) >> 32).  The host generator
 * (cadr_amd/scenes.py) uses the same streams, so a level generated here equals
 * the host's up to the last bit of the double-precision sin/cos/exp it
 * rounds to float.
 */
#ifndef CADR_SYNTH_H
#define CADR_SYNTH_H

#include "cadr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cadrSynthTransform {
	/* uniform random rotation (Shoemake quaternion from u0, u1, u2) plus
	 * translation -range + 2*range*u{3,4,5}; 6 components per item */
	CADR_SYNTH_ROTATION_TRANSLATION = 0,
	/* uniform scale exp(ln 0.1 + u0*(ln 2 - ln 0.1)) plus translation
	 * -range + 2*range*u{1,2,3}; 4 components per item */
	CADR_SYNTH_SCALE_TRANSLATION = 1
} cadrSynthTransform;

/* One level of a synthetic transform tree, or a flat scene.
 *
 * Node i of the level is tree node t(i): t(i) = i, or, with permuteKey != 0,
 * a pseudo-random permutation of [0, count) (a 4-round Feistel network on the
 * smallest even-bit domain covering count, cycle-walked into range).  A leaf
 * level ordered by StateSet uses the permutation: consecutive StateSet ranges
 * then hold uniformly random sets of tree nodes.
 *
 * Outputs (each optional, 0 = not written), indexed by i:
 *   localMatrices  float mat4, column-major, from (seed, stream, t(i))
 *   parentIndex    t(i) / fanout (uint32)
 *   meshIndex      uniform integer in [0, numMeshes-1] from (seed, meshStream, t(i))
 *   spheres        float4 world sphere of mesh meshIndex under the local
 *                  matrix (the transform of cadrPropagateTransforms for a
 *                  root level), for flat scenes; needs meshIndex and meshes */
typedef struct cadrSynthLevel {
	uint64_t seed;
	uint64_t count;
	uint64_t permuteKey;
	uint64_t localMatrices;
	uint64_t parentIndex;
	uint64_t meshIndex;
	uint64_t meshes;            /* cadrMesh[numMeshes] (device), for spheres */
	uint64_t spheres;
	uint32_t stream;            /* RNG stream of the transforms */
	uint32_t transform;         /* cadrSynthTransform */
	float    translationRange;
	uint32_t fanout;            /* must be > 0 when parentIndex != 0 */
	uint32_t meshStream;
	uint32_t numMeshes;         /* must be > 0 when meshIndex != 0 */
} cadrSynthLevel;               /* 88 B */

int cadrSynthesizeLevel(const cadrSynthLevel* level, cadrStream stream);

/* The permutation t(i) above, on the host (for checks and host-side mirrors). */
uint64_t cadrSynthPermute(uint64_t index, uint64_t count, uint64_t key);

#ifdef __cplusplus
}
#endif

#endif /* CADR_SYNTH_H */
